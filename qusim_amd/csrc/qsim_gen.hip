// qsim_gen.hip -- device-side generation of parameterised inputs: potentials
// made of Gaussian bumps plus an imaginary absorbing layer, and Gaussian wave
// packets.  The reference evaluates a std::function at every grid point on
// the host and copies the result over (src/EvolverImpl1D.cpp:40-70,
// src/EvolverImpl2D.cpp:40-99, src/Perturbation3DImpl.cpp:28-50); a request
// loop that builds thousands of grids cannot afford that, so the same
// functions are evaluated here from their parameters, point by point, in the
// order of operations of the host formulas below.
//
// Grid: x_a[i] = lo[a] + (hi[a] - lo[a]) / n[a] * i  (include/QuSim.h:47-61),
// row-major, last axis fastest.
#include <cmath>

#include "qsim_common.h"

using namespace qsim;

namespace {

constexpr int kMaxDim = 3;
constexpr int kMaxBumps = 64;

struct GridG {
    int dim;
    size_t n[kMaxDim];
    double lo[kMaxDim], hi[kMaxDim], d[kMaxDim];
    size_t total;
};

struct BumpsG {
    int count;
    // per bump: height, width, centre[dim]
    double h[kMaxBumps], w[kMaxBumps], c[kMaxBumps][kMaxDim];
    double absorb_width, absorb_strength;
};

__device__ __forceinline__ void coords(const GridG &g, size_t idx, double *x)
{
    for (int a = g.dim - 1; a >= 0; --a) {
        const size_t i = idx % g.n[a];
        idx /= g.n[a];
        x[a] = g.lo[a] + g.d[a] * (double)i;
    }
}

// V = sum_b h_b exp(-(sum_a (x_a - c_ba)^2) / (2 w_b w_b))
//     - i s ((W - d) / W)^2 for d < W, d = min_a min(x_a - lo_a, hi_a - x_a)
__global__ void __launch_bounds__(256)
    k_gen_potential(GridG g, BumpsG b, double *re, double *im, c128 *cv)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < g.total;
         idx += stride) {
        double x[kMaxDim];
        coords(g, idx, x);
        double v = 0.0;
        for (int k = 0; k < b.count; ++k) {
            double r2 = 0.0;
            for (int a = 0; a < g.dim; ++a) {
                const double t = x[a] - b.c[k][a];
                r2 = a == 0 ? t * t : r2 + t * t;
            }
            v += b.h[k] * exp(-r2 / (2.0 * b.w[k] * b.w[k]));
        }
        double w = 0.0;
        if (b.absorb_width > 0.0) {
            double dmin = 0.0;
            for (int a = 0; a < g.dim; ++a) {
                const double da = fmin(x[a] - g.lo[a], g.hi[a] - x[a]);
                dmin = a == 0 ? da : fmin(dmin, da);
            }
            if (dmin < b.absorb_width) {
                const double q = (b.absorb_width - dmin) / b.absorb_width;
                w = b.absorb_strength * (q * q);
            }
        }
        if (re) re[idx] = v;
        if (im) im[idx] = -w;
        if (cv) cv[idx] = mk(v, -w);
    }
}

// psi = exp(i k.x - (sum_a (x_a - c_a)^2) / (4 sigma^2))
__global__ void __launch_bounds__(256)
    k_gen_packet(GridG g, double c0, double c1, double c2, double k0, double k1, double k2,
                 double sigma, c128 *psi)
{
    const double c[3] = {c0, c1, c2}, kk[3] = {k0, k1, k2};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < g.total;
         idx += stride) {
        double x[kMaxDim];
        coords(g, idx, x);
        double ph = 0.0, r2 = 0.0;
        for (int a = 0; a < g.dim; ++a) {
            const double t = x[a] - c[a];
            ph = a == 0 ? kk[a] * x[a] : ph + kk[a] * x[a];
            r2 = a == 0 ? t * t : r2 + t * t;
        }
        const double amp = exp(-r2 / (4.0 * (sigma * sigma)));
        double sn, cs;
        sincos(ph, &sn, &cs);
        psi[idx] = mk(amp * cs, amp * sn);
    }
}

int make_grid(GridG *g, int dim, const size_t *n, const double *lo, const double *hi)
{
    if (dim < 1 || dim > kMaxDim || !n || !lo || !hi)
        return set_error(QSIM_ERR_INVALID, "gen: dim must be 1..3 with n, lo, hi given");
    g->dim = dim;
    g->total = 1;
    for (int a = 0; a < kMaxDim; ++a) {
        g->n[a] = 1;
        g->lo[a] = g->hi[a] = g->d[a] = 0.0;
    }
    for (int a = 0; a < dim; ++a) {
        if (n[a] == 0 || !(hi[a] > lo[a]))
            return set_error(QSIM_ERR_INVALID, "gen: empty axis or hi <= lo");
        g->n[a] = n[a];
        g->lo[a] = lo[a];
        g->hi[a] = hi[a];
        g->d[a] = (hi[a] - lo[a]) / (double)n[a];
        g->total *= n[a];
    }
    return QSIM_OK;
}

unsigned blocks_for(size_t n)
{
    size_t b = (n + 255) / 256;
    const size_t cap = (size_t)device_cu_count() * 16;
    return (unsigned)(b < cap ? (b ? b : 1) : cap);
}

}  // namespace

extern "C" {

int qsim_hip_gen_potential(int dim, const size_t *n, const double *lo, const double *hi,
                           const double *bumps, int count, double absorb_width,
                           double absorb_strength, double *re, double *im, qsim_c128 *cv,
                           qsim_stream_t s)
{
    GridG g;
    QSIM_TRY(make_grid(&g, dim, n, lo, hi));
    if (count < 0 || count > kMaxBumps || (count > 0 && !bumps))
        return set_error(QSIM_ERR_INVALID, "gen_potential: 0..64 bumps, parameters on the host");
    if (!re && !im && !cv) return set_error(QSIM_ERR_INVALID, "gen_potential: no output");
    BumpsG b;
    b.count = count;
    const int per = 2 + dim;
    for (int k = 0; k < count; ++k) {
        b.h[k] = bumps[k * per + 0];
        b.w[k] = bumps[k * per + 1];
        if (!(b.w[k] > 0.0)) return set_error(QSIM_ERR_INVALID, "gen_potential: width <= 0");
        for (int a = 0; a < kMaxDim; ++a) b.c[k][a] = a < dim ? bumps[k * per + 2 + a] : 0.0;
    }
    b.absorb_width = absorb_width;
    b.absorb_strength = absorb_strength;
    hipLaunchKernelGGL(k_gen_potential, dim3(blocks_for(g.total)), dim3(256), 0, as_stream(s), g,
                       b, re, im, reinterpret_cast<c128 *>(cv));
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

int qsim_hip_norm2(const qsim_c128 *in, size_t n, double *out, qsim_stream_t s);
int qsim_hip_scale(qsim_c128 *fromto, double scale, size_t n, qsim_stream_t s);

int qsim_hip_gen_packet(int dim, const size_t *n, const double *lo, const double *hi,
                        const double *center, const double *k, double sigma, int normalise,
                        qsim_c128 *psi, qsim_stream_t s)
{
    GridG g;
    QSIM_TRY(make_grid(&g, dim, n, lo, hi));
    if (!center || !k || !psi || !(sigma > 0.0))
        return set_error(QSIM_ERR_INVALID, "gen_packet: centre, k, psi and sigma > 0 needed");
    double c[3] = {0, 0, 0}, kk[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        c[a] = center[a];
        kk[a] = k[a];
    }
    hipLaunchKernelGGL(k_gen_packet, dim3(blocks_for(g.total)), dim3(256), 0, as_stream(s), g,
                       c[0], c[1], c[2], kk[0], kk[1], kk[2], sigma, reinterpret_cast<c128 *>(psi));
    QSIM_HIP(hipGetLastError());
    if (normalise) {  // sum |psi|^2 dV = 1, as force_normalization (src/EvolverImpl2D.cpp:87-92)
        double n2 = 0.0;
        QSIM_TRY(qsim_hip_norm2(psi, g.total, &n2, s));
        double dv = 1.0;
        for (int a = 0; a < dim; ++a) dv *= g.d[a];
        if (!(n2 > 0.0)) return set_error(QSIM_ERR_INVALID, "gen_packet: zero norm");
        QSIM_TRY(qsim_hip_scale(psi, 1.0 / std::sqrt(n2 * dv), g.total, s));
    }
    return QSIM_OK;
}

}  // extern "C"
