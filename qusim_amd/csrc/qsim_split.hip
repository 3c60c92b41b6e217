// qsim_split.hip -- split-operator evolution job (SplittingMethod1D/2D on
// gfx950).
//
// Reference step (src/SplittingUtils.h:14-64, src/SplittingMethod2D.cpp:81-121):
//   psi <- E_V psi ; psi <- IFFT(E_T FFT(psi)) / N ; psi <- E_V psi
// Fused schedule (2D, power-of-two axes): the potential half steps of
// consecutive steps meet in real space, so with FFT2 = FFT_x FFT_y a run of
// steps becomes alternating sweeps, each one read + one write of psi:
//   row sweep  (y lines, contiguous): IFFT_y -> x E_V(x,y) -> FFT_y
//   col sweep  (x lines, stride ny):  FFT_x  -> x Tx(kx) Ty(ky) -> IFFT_x
// with E_T = Tx Ty separable and 1/(nx ny) folded into Tx.  The first sweep
// only applies E_V then FFT_y; the last only IFFT_y then E_V.
// 1D (power-of-two): one workgroup per job keeps psi in registers for all
// steps.  Other sizes and the infinite-wall boundary run the reference's op
// sequence unfused (FFT plans + pointwise kernels).
//
// Precision 64: complex128 everywhere.  Precision 32 (FP32-mixed): psi is
// stored and transformed in complex64; every phase factor and inter-pass
// twiddle is FP64 and its product with the data is formed in FP64 and rounded
// once, so the rounding of a factor is never repeated step after step.
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "qsim_fft_kernels.h"
#include "qsim_fft_plan.h"

namespace qsim {

enum { SW_FIRST = 0, SW_MID = 1, SW_LAST = 2 };

// Device-side description of one job (the kernels index a table of these by
// blockIdx.y, so many same-shape jobs share one launch).
struct SplitDev {
    c128 *psi;             // caller's complex128 state
    void *work;            // working state (== psi for precision 64)
    const c128 *ev[4];     // potential factors over the grid
    const c128 *kx[2];     // kinetic factor along x (2D) or the whole line (1D)
    const c128 *ky[2];     // kinetic factor along y (2D)
};

// v * f with f an FP64 factor.  FP32 data is widened, multiplied in FP64 and
// rounded once, so a near-identity factor (tiny V dt) keeps its sub-ulp part
// instead of losing it the same way at every step.  A float-float product
// that rounded v * hi before adding v * lo lost it: at 2048^2, dt = 0.1, that
// bias was half of a 1.1e-5 psi error after 1000 steps.
__device__ __forceinline__ c128 apply(c128 v, c128 f) { return v * f; }
__device__ __forceinline__ c64 apply(c64 v, c128 f) { return cvt<float>(cvt<double>(v) * f); }

template <typename TO, typename TI>
__device__ __forceinline__ cx<TO> conv(cx<TI> a)
{
    return cvt<TO>(a);
}

// Sweep input lines are read once per sweep.  A non-temporal load keeps them
// from displacing E_V, the kinetic tables and the twiddles in L2, which the
// next step reads again.  Measured per launch: with at least 8 MiB of lines
// of up to 1024 points it pays (1024^2 FP64 4.22 -> 4.45e10 grid-pt-steps/s,
// FP32 4.73 -> 5.26e10, 16 batched 256^2 FP32 jobs 5.2 -> 6.4e10); smaller
// launches lose 3-7% (their lines stay in L2 anyway) and 2048-point lines
// lose 10% at 2048^2 FP64.  The sweeps take the form as a template argument
// (a runtime branch around the loads lost the gain) and the host picks it
// per launch (line_nt).
template <bool NT, typename T>
__device__ __forceinline__ cx<T> ld_line(const cx<T> *p)
{
    if constexpr (NT) return ld_nt(p);
    else return ld_g(p);
}
static bool line_nt(int n, size_t line_bytes_per_launch)
{
    return n <= 1024 && line_bytes_per_launch >= ((size_t)8 << 20);
}

// ---------------------------------------------------------------- 2D sweeps
// Row sweep over the L rows g0 .. g0 + L - 1 of one job (Line P, L lines of
// P::TPL threads, LDS line stride LSTR): the body of k_row_sweep and of the
// multi-size group kernel.
template <typename T, typename P, int L, int LSTR, int MODE, bool NT>
__device__ __forceinline__ void row_lines(const SplitDev &J, size_t g0, size_t nrows, int iev,
                                          const c128 *__restrict__ tw, cx<T> *lds)
{
    constexpr int N = P::TPL * P::E;
    const int tid = threadIdx.x;
    const int l = tid / P::TPL, t = tid % P::TPL;
    const size_t g = g0 + l;
    const bool active = g < nrows;
    const size_t base = (active ? g : 0) * N;
    const c128 *__restrict__ ev = J.ev[iev];
    cx<T> *work = reinterpret_cast<cx<T> *>(J.work);
    c128 *psi = J.psi;
    constexpr int RA = P::RA_, RZ = P::RLAST;
    cx<T> v[P::E];
    if constexpr (MODE == SW_FIRST) {
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r) {
                const int pos = P::template in_pos<RA>(t, u, r);
                v[u * RA + r] =
                    active ? apply(conv<T>(ld_g(psi + base + pos)), ld_g(ev + base + pos))
                           : mk<T>(0, 0);
            }
    } else {
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                v[u * RZ + r] =
                    active ? ld_line<NT>(work + base + P::template in_pos<RZ>(t, u, r)) : mk<T>(0, 0);
        // the potential factors do not depend on the transform: issue their
        // loads first so they land while the inverse FFT runs
        constexpr bool PREFETCH = P::E <= 16;
        c128 evp[PREFETCH ? P::E : 1];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r)
                    evp[u * RA + r] = active ? ld_g(ev + base + P::template in_pos<RA>(t, u, r))
                                             : mk(0.0, 0.0);
        }
        P::inverse(v, t, lds + l * LSTR, tw);
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r) {
                const int pos = P::template in_pos<RA>(t, u, r);
                if constexpr (PREFETCH) {
                    v[u * RA + r] = apply(v[u * RA + r], evp[u * RA + r]);
                } else {
                    if (active) v[u * RA + r] = apply(v[u * RA + r], ld_g(ev + base + pos));
                }
            }
    }
    if constexpr (MODE == SW_LAST) {
        if (active) {
#pragma unroll
            for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r)
                    store_nt(&psi[base + P::template in_pos<RA>(t, u, r)],
                             conv<double>(v[u * RA + r]));
        }
    } else {
        P::forward(v, t, lds + l * LSTR, tw);
        if (active) {
#pragma unroll
            for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    store_nt(&work[base + P::template in_pos<RZ>(t, u, r)], v[u * RZ + r]);
        }
    }
}

template <typename T, int N, int MODE, int V, bool NT>
__global__ void __launch_bounds__((P2Launch<T, N, false, V>::THREADS))
    k_row_sweep(const SplitDev *__restrict__ jobs, size_t nrows, int iev,
                const c128 *__restrict__ tw)
{
    using LP = P2Launch<T, N, false, V>;
    extern __shared__ unsigned char smem_raw[];
    row_lines<T, typename LP::Line, LP::L, LP::LSTR, MODE, NT>(
        jobs[blockIdx.y], (size_t)blockIdx.x * LP::L, nrows, iev, tw,
        reinterpret_cast<cx<T> *>(smem_raw));
}

// Column sweep over the L columns c0 .. c0 + L - 1 of one job's nx(=N) x
// ncols array, lanes line-fast (lane i on column i % L).
template <typename T, typename P, int L, int LSTR, bool NT>
__device__ __forceinline__ void col_lines(const SplitDev &J, size_t c0, size_t ncols, int ik,
                                          const c128 *__restrict__ tw, cx<T> *lds)
{
    const int tid = threadIdx.x;
    const int l = tid % L, t = tid / L;
    const size_t j = c0 + l;
    const bool active = j < ncols;
    const size_t c = active ? j : 0;
    cx<T> *work = reinterpret_cast<cx<T> *>(J.work);
    const c128 *__restrict__ kx = J.kx[ik];
    const c128 kyj = ld_g(J.ky[ik] + c);
    constexpr int RA = P::RA_, RZ = P::RLAST;
    cx<T> v[P::E];
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r)
            v[u * RA + r] =
                active ? ld_line<NT>(work + c + (size_t)P::template in_pos<RA>(t, u, r) * ncols)
                       : mk<T>(0, 0);
    // the kinetic factors depend only on the thread's output positions:
    // issue their loads with the data so they land during the forward FFT
    constexpr bool PREFETCH = P::E <= 8;
    c128 kxp[PREFETCH ? P::E : 1];
    if constexpr (PREFETCH) {
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                kxp[u * RZ + r] = ld_g(kx + P::template in_pos<RZ>(t, u, r));
    }
    P::forward(v, t, lds + l * LSTR, tw);
#pragma unroll
    for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const c128 kk = PREFETCH ? kxp[u * RZ + r] : ld_g(kx + P::template in_pos<RZ>(t, u, r));
            v[u * RZ + r] = apply(v[u * RZ + r], kk * kyj);
        }
    P::inverse(v, t, lds + l * LSTR, tw);
    if (active) {
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r)
                store_nt(&work[c + (size_t)P::template in_pos<RA>(t, u, r) * ncols], v[u * RA + r]);
    }
}

template <typename T, int N, int V, bool NT>
__global__ void __launch_bounds__((P2Launch<T, N, true, V>::THREADS))
    k_col_sweep(const SplitDev *__restrict__ jobs, size_t ncols, int ik,
                const c128 *__restrict__ tw)
{
    using LP = P2Launch<T, N, true, V>;
    extern __shared__ unsigned char smem_raw[];
    col_lines<T, typename LP::Line, LP::L, LP::LSTR, NT>(
        jobs[blockIdx.y], xcd_tile(blockIdx.x, gridDim.x) * LP::L, ncols, ik, tw,
        reinterpret_cast<cx<T> *>(smem_raw));
}

// ---------------------------------------------------------------- 1D job
// Phase program: start with ev[ev0]; each step runs (kinetic k[i], potential
// ev[e[i]]) for i < nph; the last potential factor of the last step is
// ev[ev_end] instead.
struct Prog1D {
    int ev0, ev_end, nph;
    int k[3], e[3];
};

// Radix plan of the persistent 1D kernel per length.  A job is one workgroup
// stepping its line on its own, so each step's chain of dependent work sets
// the time: lengths 1024, 2048 and 4096 take four-pass plans with 8 elements
// per thread (128 / 256 / 512 threads) instead of the 2D defaults (32 / 128 /
// 256 threads).
template <typename T, int N>
struct Split1DV {
    static constexpr int V = (N == 1024 || N == 2048) ? 6 : (N == 4096 ? 1 : 0);
};

template <typename T, int N>
__global__ void __launch_bounds__((P2Cfg<T, N, Split1DV<T, N>::V>::Line::TPL))
    k_split1d(const SplitDev *__restrict__ jobs, Prog1D prog, size_t nsteps,
              const c128 *__restrict__ tw)
{
    using P = typename P2Cfg<T, N, Split1DV<T, N>::V>::Line;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const SplitDev &J = jobs[blockIdx.x];
    const int t = threadIdx.x;
    constexpr int RA = P::RA_, RZ = P::RLAST;
    // with 8 elements per thread the step's factors fit in registers: their
    // loads are issued at the top of the step and land during the forward FFT
    constexpr bool PREFETCH = P::E <= 8;
    cx<T> v[P::E];
    c128 *psi = J.psi;
    const c128 *__restrict__ e0 = J.ev[prog.ev0];
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int pos = P::template in_pos<RA>(t, u, r);
            v[u * RA + r] = apply(conv<T>(ld_g(psi + pos)), ld_g(e0 + pos));
        }
    for (size_t s = 0; s < nsteps; ++s) {
        for (int ph = 0; ph < prog.nph; ++ph) {
            // the twiddle loads do not change from step to step; an opaque
            // pointer keeps the compiler from hoisting every power of every
            // pass out of the step loop into registers (that spilled)
            const c128 *twl = tw;
            __asm__ volatile("" : "+s"(twl));
            const bool last = (s + 1 == nsteps) && (ph + 1 == prog.nph);
            if constexpr (PREFETCH) {
                const c128 *kt = J.kx[prog.k[ph]];
                const c128 *ev = J.ev[last ? prog.ev_end : prog.e[ph]];
                c128 kp[P::E], ep[P::E];
#pragma unroll
                for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
                    for (int r = 0; r < RZ; ++r)
                        kp[u * RZ + r] = ld_g(kt + P::template in_pos<RZ>(t, u, r));
#pragma unroll
                for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                    for (int r = 0; r < RA; ++r)
                        ep[u * RA + r] = ld_g(ev + P::template in_pos<RA>(t, u, r));
                P::forward(v, t, lds, twl);
#pragma unroll
                for (int i = 0; i < P::E; ++i) v[i] = apply(v[i], kp[i]);
                P::inverse(v, t, lds, twl);
#pragma unroll
                for (int i = 0; i < P::E; ++i) v[i] = apply(v[i], ep[i]);
            } else {
                P::forward(v, t, lds, twl);
                const c128 *kt = J.kx[prog.k[ph]];
                __asm__ volatile("" : "+s"(kt));  // likewise: reload each step
#pragma unroll
                for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
                    for (int r = 0; r < RZ; ++r) {
                        const int k = P::template in_pos<RZ>(t, u, r);
                        v[u * RZ + r] = apply(v[u * RZ + r], ld_g(kt + k));
                    }
                P::inverse(v, t, lds, twl);
                const c128 *ev = J.ev[last ? prog.ev_end : prog.e[ph]];
                __asm__ volatile("" : "+s"(ev));
#pragma unroll
                for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                    for (int r = 0; r < RA; ++r) {
                        const int pos = P::template in_pos<RA>(t, u, r);
                        v[u * RA + r] = apply(v[u * RA + r], ld_g(ev + pos));
                    }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r)
            st_g(psi + P::template in_pos<RA>(t, u, r), conv<double>(v[u * RA + r]));
}

// ---------------------------------------------------------------- helpers
template <typename F>
__global__ void __launch_bounds__(256) k_each(size_t n, F f)
{
    size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) f(i);
}

template <typename F>
int each(size_t n, hipStream_t s, F f)
{
    if (!n) return QSIM_OK;
    size_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_each<F>, dim3((unsigned)b), dim3(256), 0, s, n, f);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

using cd = std::complex<double>;

}  // namespace qsim

using namespace qsim;

// captured step sequences kept per job (qsim_hip_split_use_graph)
constexpr size_t kMaxGraphs = 8;

struct qsim_split_job {
    qsim_split_desc d{};
    size_t n = 0;
    bool fused = false;
    c128 *psi = nullptr;
    void *work = nullptr;   // owned when precision 32
    std::vector<void *> owned;
    SplitDev host{};
    SplitDev *dev = nullptr;  // device copy of `host`
    const void *tw_x = nullptr, *tw_y = nullptr;  // twiddles of the sweep kernels
    // launch variants of the row / column kernels (autotuned or default 0)
    int row_var = 0, col_var = 0;
    // replay captured sweep sequences (keyed by step count, oldest first in
    // graph_order)
    bool use_graph = false;
    std::map<size_t, hipGraphExec_t> graphs;
    std::vector<size_t> graph_order;
    // phase program
    int ev_start = 0, ev_end = 0, nph = 0;
    int ph_k[3] = {0, 0, 0}, ph_e[3] = {0, 0, 0};
    // unfused path
    qsim_fft_plan *fwd = nullptr, *inv = nullptr;
    c128 *tmp = nullptr, *tmp2 = nullptr;
    c128 *et_full[2] = {nullptr, nullptr};
};

namespace {

// O4 coefficients, src/SplittingUtils.cpp:3-13
struct O4 {
    double C1, C2, D1, D2;
    O4()
    {
        double t = std::pow(2, 1 / 3.0);
        C1 = 1 / (2 * (2 - t));
        C2 = (1 - t) / (2 * (2 - t));
        D1 = 1 / (2 - t);
        D2 = -t / (2 - t);
    }
};

int dalloc(qsim_split_job *j, void **p, size_t bytes)
{
    QSIM_HIP(hipMalloc(p, bytes));
    j->owned.push_back(*p);
    return QSIM_OK;
}

// ev = exp(f V), f = -I/hbar * dt * t (src/SplittingMethod2D.cpp:43-52)
int build_ev(qsim_split_job *j, c128 **out, double t, const double *Vre, const double *Vim,
             hipStream_t s)
{
    void *p;
    QSIM_TRY(dalloc(j, &p, j->n * sizeof(c128)));
    c128 *ev = (c128 *)p;
    cd f = cd(-0.0, -1.0) / j->d.hbar * cd(j->d.dt_re, j->d.dt_im) * t;
    c128 fc = mk(f.real(), f.imag());
    QSIM_TRY(each(j->n, s, [=] __device__(size_t i) {
        c128 z = Vim ? fc * mk(Vre[i], Vim[i]) : fc * Vre[i];
        ev[i] = cexp(z);
    }));
    *out = ev;
    return QSIM_OK;
}

int square_ev(qsim_split_job *j, c128 **out, const c128 *a, hipStream_t s)
{
    void *p;
    QSIM_TRY(dalloc(j, &p, j->n * sizeof(c128)));
    c128 *ev = (c128 *)p;
    QSIM_TRY(each(j->n, s, [=] __device__(size_t i) { ev[i] = a[i] * a[i]; }));
    *out = ev;
    return QSIM_OK;
}

// exp(alpha * a * fold(i)^2) * scale over one axis of length n (host, FP64)
int build_axis(qsim_split_job *j, c128 **out, size_t n, cd alpha, double a, double scale)
{
    std::vector<c128> h(n);
    for (size_t i = 0; i < n; ++i) {
        double k = (double)fold_half(i, n);
        cd e = std::exp(alpha * (k * k * a)) * scale;
        h[i] = mk(e.real(), e.imag());
    }
    void *p;
    QSIM_TRY(dalloc(j, &p, n * sizeof(c128)));
    QSIM_HIP(hipMemcpy(p, h.data(), n * sizeof(c128), hipMemcpyHostToDevice));
    *out = (c128 *)p;
    return QSIM_OK;
}

template <typename T, int N, int MODE, int V>
int launch_row(const SplitDev *jobs, int njobs, size_t nrows, int iev, const void *tw,
               hipStream_t s)
{
    using LP = P2Launch<T, N, false, V>;
    auto fn = k_row_sweep<T, N, MODE, V, false>;
    if constexpr (N <= 1024)
        if (line_nt(N, (size_t)njobs * nrows * N * sizeof(cx<T>))) fn = k_row_sweep<T, N, MODE, V, true>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    dim3 grid((unsigned)((nrows + LP::L - 1) / LP::L), (unsigned)njobs);
    hipLaunchKernelGGL(fn, grid, dim3(LP::THREADS), LP::LDS_BYTES, s, jobs, nrows, iev,
                       (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

template <typename T, int N, int V>
int launch_col(const SplitDev *jobs, int njobs, size_t ncols, int ik, const void *tw,
               hipStream_t s)
{
    using LP = P2Launch<T, N, true, V>;
    auto fn = k_col_sweep<T, N, V, false>;
    if constexpr (N <= 1024)
        if (line_nt(N, (size_t)njobs * ncols * N * sizeof(cx<T>))) fn = k_col_sweep<T, N, V, true>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    dim3 grid((unsigned)((ncols + LP::L - 1) / LP::L), (unsigned)njobs);
    hipLaunchKernelGGL(fn, grid, dim3(LP::THREADS), LP::LDS_BYTES, s, jobs, ncols, ik,
                       (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

template <typename T, int N>
int launch_1d(const SplitDev *jobs, int njobs, const Prog1D &prog, size_t nsteps, const void *tw,
              hipStream_t s)
{
    using P = typename P2Cfg<T, N, Split1DV<T, N>::V>::Line;
    auto fn = k_split1d<T, N>;
    const size_t lds = (size_t)P::LS * sizeof(cx<T>);
    QSIM_TRY(lds_limit((const void *)fn, lds));
    hipLaunchKernelGGL(fn, dim3((unsigned)njobs), dim3(P::TPL), lds, s, jobs, prog, nsteps,
                       (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

// dispatch on (precision, length); returns UNSUPPORTED for lengths without a
// fused instance
#define QSIM_SWITCH_N(PREC, NVAL, CALL)                                               \
    switch (NVAL) {                                                                   \
        case 64: return PREC == 64 ? CALL(double, 64) : CALL(float, 64);              \
        case 128: return PREC == 64 ? CALL(double, 128) : CALL(float, 128);           \
        case 256: return PREC == 64 ? CALL(double, 256) : CALL(float, 256);           \
        case 512: return PREC == 64 ? CALL(double, 512) : CALL(float, 512);           \
        case 1024: return PREC == 64 ? CALL(double, 1024) : CALL(float, 1024);        \
        case 2048: return PREC == 64 ? CALL(double, 2048) : CALL(float, 2048);        \
        case 4096: return PREC == 64 ? CALL(double, 4096) : CALL(float, 4096);        \
        case 8192:                                                                    \
            if (PREC == 32) return CALL(float, 8192);                                 \
            break;                                                                    \
        default: break;                                                               \
    }                                                                                 \
    return set_error(QSIM_ERR_UNSUPPORTED, "no fused kernel for this length");

// Launch variants: lengths 1024 and 2048 have the candidates 0..kVariants-1
// (P2Cfg), every other length only variant 0.
constexpr int kVariants = 8;

bool has_variants(size_t n) { return n == 1024 || n == 2048; }

// Variants a new job starts with: the pairs the autotuner picked most often
// on MI355X for single and batched jobs of these lengths (row variant by the
// row length ny, column variant by the column length nx).  Fixed, so jobs that
// are never tuned (request loops, batches) still take a fast pair and two
// runs of one job always choose the same radix plans.
int default_variant(int prec, size_t n, bool row)
{
    if (n == 1024) return prec == 64 ? 4 : 7;
    if (n == 2048) return prec == 64 ? (row ? 7 : 5) : (row ? 7 : 3);
    return 0;
}

template <typename T, int NN, int MODE>
int row_v(int var, const SplitDev *jobs, int njobs, size_t nrows, int iev, const void *tw,
          hipStream_t s)
{
    if constexpr (NN == 1024 || NN == 2048) {
        switch (var) {
            case 1: return launch_row<T, NN, MODE, 1>(jobs, njobs, nrows, iev, tw, s);
            case 2: return launch_row<T, NN, MODE, 2>(jobs, njobs, nrows, iev, tw, s);
            case 3: return launch_row<T, NN, MODE, 3>(jobs, njobs, nrows, iev, tw, s);
            case 4: return launch_row<T, NN, MODE, 4>(jobs, njobs, nrows, iev, tw, s);
            case 5: return launch_row<T, NN, MODE, 5>(jobs, njobs, nrows, iev, tw, s);
            case 6: return launch_row<T, NN, MODE, 6>(jobs, njobs, nrows, iev, tw, s);
            case 7: return launch_row<T, NN, MODE, 7>(jobs, njobs, nrows, iev, tw, s);
            default: break;
        }
    }
    return launch_row<T, NN, MODE, 0>(jobs, njobs, nrows, iev, tw, s);
}

template <typename T, int NN>
int col_v(int var, const SplitDev *jobs, int njobs, size_t ncols, int ik, const void *tw,
          hipStream_t s)
{
    if constexpr (NN == 1024 || NN == 2048) {
        switch (var) {
            case 1: return launch_col<T, NN, 1>(jobs, njobs, ncols, ik, tw, s);
            case 2: return launch_col<T, NN, 2>(jobs, njobs, ncols, ik, tw, s);
            case 3: return launch_col<T, NN, 3>(jobs, njobs, ncols, ik, tw, s);
            case 4: return launch_col<T, NN, 4>(jobs, njobs, ncols, ik, tw, s);
            case 5: return launch_col<T, NN, 5>(jobs, njobs, ncols, ik, tw, s);
            case 6: return launch_col<T, NN, 6>(jobs, njobs, ncols, ik, tw, s);
            case 7: return launch_col<T, NN, 7>(jobs, njobs, ncols, ik, tw, s);
            default: break;
        }
    }
    return launch_col<T, NN, 0>(jobs, njobs, ncols, ik, tw, s);
}

int row_sweep(int prec, size_t ny, int mode, int var, const SplitDev *jobs, int njobs,
              size_t nrows, int iev, const void *tw, hipStream_t s)
{
#define ROWCALL(T, NN)                                                                     \
    (mode == SW_FIRST ? row_v<T, NN, SW_FIRST>(var, jobs, njobs, nrows, iev, tw, s)         \
                      : mode == SW_MID ? row_v<T, NN, SW_MID>(var, jobs, njobs, nrows, iev, tw, s) \
                                       : row_v<T, NN, SW_LAST>(var, jobs, njobs, nrows, iev, tw, s))
    QSIM_SWITCH_N(prec, ny, ROWCALL)
#undef ROWCALL
}

int col_sweep(int prec, size_t nx, int var, const SplitDev *jobs, int njobs, size_t ncols,
              int ik, const void *tw, hipStream_t s)
{
#define COLCALL(T, NN) col_v<T, NN>(var, jobs, njobs, ncols, ik, tw, s)
    QSIM_SWITCH_N(prec, nx, COLCALL)
#undef COLCALL
}

int run_1d(int prec, size_t n, const SplitDev *jobs, int njobs, const Prog1D &prog,
           size_t nsteps, const void *tw, hipStream_t s)
{
    // one workgroup holds the whole line: 8192 (FP64) and 16384 (FP32) fill
    // 139 KB of LDS, so they exist for the persistent 1D kernel only
    if (prec == 64 && n == 8192) return launch_1d<double, 8192>(jobs, njobs, prog, nsteps, tw, s);
    if (prec == 32 && n == 16384) return launch_1d<float, 16384>(jobs, njobs, prog, nsteps, tw, s);
#define ONECALL(T, NN) launch_1d<T, NN>(jobs, njobs, prog, nsteps, tw, s)
    QSIM_SWITCH_N(prec, n, ONECALL)
#undef ONECALL
}

// lengths with a fused kernel: 2D axes 64..4096 (FP64) / 8192 (FP32); 1D
// lines up to twice that
bool fused_len_ok(size_t n, int prec, int dim)
{
    if (n < 64 || (n & (n - 1))) return false;
    const size_t cap = (prec == 64 ? 4096 : 8192) * (dim == 1 ? 2 : 1);
    return n <= cap;
}

// Optional per-kernel timing: HIP events around every sweep launch.
enum { TK_ROW = 0, TK_COL = 1, TK_1D = 2, TK_ROW_FIRST = 3, TK_ROW_LAST = 4, TK_KINDS = 5 };
struct SweepTimer {
    std::vector<hipEvent_t> ev;  // pairs (start, stop)
    std::vector<int> kind;       // a TK_* per pair
    int mark(int k, hipStream_t s)
    {
        hipEvent_t e;
        // timing-only events: no system-scope fence (cache writeback and
        // invalidate) between the kernels they bracket
        QSIM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        ev.push_back(e);
        if (k >= 0) kind.push_back(k);
        QSIM_HIP(hipEventRecord(e, s));
        return QSIM_OK;
    }
    ~SweepTimer()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

#define TIMED(T, K, CALL)                          \
    do {                                           \
        if (T) QSIM_TRY((T)->mark((K), s));        \
        QSIM_TRY(CALL);                            \
        if (T) QSIM_TRY((T)->mark(-1, s));         \
    } while (0)

// ------------------------------------------------- fused schedule (2D / 1D)
int fused_steps(qsim_split_job *j, const SplitDev *jobs, int njobs, size_t nsteps,
                hipStream_t s, SweepTimer *tm = nullptr)
{
    const int prec = j->d.precision;
    if (j->d.dim == 1) {
        Prog1D p;
        p.ev0 = j->ev_start;
        p.ev_end = j->ev_end;
        p.nph = j->nph;
        for (int i = 0; i < 3; ++i) {
            p.k[i] = j->ph_k[i];
            p.e[i] = j->ph_e[i];
        }
        TIMED(tm, TK_1D, run_1d(prec, j->d.nx, jobs, njobs, p, nsteps, j->tw_x, s));
        return QSIM_OK;
    }
    const size_t nx = j->d.nx, ny = j->d.ny;
    const int rv = j->row_var, cv = j->col_var;
    TIMED(tm, TK_ROW_FIRST,
          row_sweep(prec, ny, SW_FIRST, rv, jobs, njobs, nx, j->ev_start, j->tw_y, s));
    for (size_t st = 0; st < nsteps; ++st) {
        for (int ph = 0; ph < j->nph; ++ph) {
            TIMED(tm, TK_COL, col_sweep(prec, nx, cv, jobs, njobs, ny, j->ph_k[ph], j->tw_x, s));
            const bool last = (st + 1 == nsteps) && (ph + 1 == j->nph);
            if (last)
                TIMED(tm, TK_ROW_LAST,
                      row_sweep(prec, ny, SW_LAST, rv, jobs, njobs, nx, j->ev_end, j->tw_y, s));
            else
                TIMED(tm, TK_ROW,
                      row_sweep(prec, ny, SW_MID, rv, jobs, njobs, nx, j->ph_e[ph], j->tw_y, s));
        }
    }
    return QSIM_OK;
}

// ------------------------------------------------- unfused reference order
// ExpV: psi *= ev (src/SplittingMethod2D.cpp:81-94); ExpT: FFT, *= E_T, IFFT,
// *1/N (:96-121).  1D InfiniteWall: src/SplittingMethod1D.cpp:62-81.
int unfused_expv(qsim_split_job *j, const c128 *ev, hipStream_t s)
{
    c128 *psi = j->psi;
    return each(j->n, s, [=] __device__(size_t i) { psi[i] = ev[i] * psi[i]; });
}

int unfused_expt(qsim_split_job *j, int ik, hipStream_t s)
{
    const size_t n = j->n;
    c128 *psi = j->psi, *ft = j->tmp;
    if (j->d.boundary == 0) {
        QSIM_TRY(qsim_hip_fft_exec(j->fwd, psi, ft, s));
        const c128 *et = j->et_full[ik];
        QSIM_TRY(each(n, s, [=] __device__(size_t i) { ft[i] = et[i] * ft[i]; }));
        QSIM_TRY(qsim_hip_fft_exec(j->inv, ft, psi, s));
        const double sc = 1. / (double)n;
        return each(n, s, [=] __device__(size_t i) { psi[i] = sc * psi[i]; });
    }
    // infinite wall: odd extension to 2N, sine transform by the 2N inverse DFT
    const size_t N = n;
    c128 *a = j->tmp, *b = j->tmp2;
    QSIM_TRY(each(2 * N, s, [=] __device__(size_t i) {
        c128 z;
        if (i == 0 || i == N) z = mk(0.0, 0.0);
        else if (i < N) z = psi[i];
        else z = mk(-psi[2 * N - i].x, -psi[2 * N - i].y);
        a[i] = z;
    }));
    QSIM_TRY(qsim_hip_fft_exec(j->inv, a, b, s));
    const c128 *et = j->et_full[ik];  // exp(f i^2), i < N
    QSIM_TRY(each(N, s, [=] __device__(size_t i) { b[i] = b[i] * et[i]; }));
    QSIM_TRY(each(2 * N, s, [=] __device__(size_t i) {
        if (i == 0 || i == N) b[i] = mk(0.0, 0.0);
        else if (i > N) b[i] = mk(-b[2 * N - i].x, -b[2 * N - i].y);
    }));
    QSIM_TRY(qsim_hip_fft_exec(j->inv, b, a, s));
    const double sc = -1. / (2. * N);
    return each(N, s, [=] __device__(size_t i) { psi[i] = sc * a[i]; });
}

int unfused_steps(qsim_split_job *j, size_t nsteps, hipStream_t s)
{
    const c128 *const *ev = j->host.ev;
    const bool o4 = j->d.method == 4;
    for (size_t st = 0; st < nsteps; ++st) {
        if (!o4) {
            QSIM_TRY(unfused_expv(j, ev[0], s));
            QSIM_TRY(unfused_expt(j, 0, s));
            QSIM_TRY(unfused_expv(j, ev[0], s));
        } else {
            QSIM_TRY(unfused_expv(j, ev[0], s));
            QSIM_TRY(unfused_expt(j, 0, s));
            QSIM_TRY(unfused_expv(j, ev[1], s));
            QSIM_TRY(unfused_expt(j, 1, s));
            QSIM_TRY(unfused_expv(j, ev[1], s));
            QSIM_TRY(unfused_expt(j, 0, s));
            QSIM_TRY(unfused_expv(j, ev[0], s));
        }
    }
    return QSIM_OK;
}

// Captured graphs may still be replaying on some stream: wait for the device
// before destroying them (only when the job has any, so jobs that never use
// graphs never wait here).
void drop_graphs(qsim_split_job *j)
{
    if (j->graphs.empty()) return;
    (void)hipDeviceSynchronize();
    for (auto &kv : j->graphs) (void)hipGraphExecDestroy(kv.second);
    j->graphs.clear();
    j->graph_order.clear();
}

void destroy(qsim_split_job *j)
{
    if (!j) return;
    drop_graphs(j);
    for (void *p : j->owned) (void)hipFree(p);
    if (j->fwd) qsim_hip_fft_plan_destroy(j->fwd);
    if (j->inv) qsim_hip_fft_plan_destroy(j->inv);
    delete j;
}

int create(qsim_split_job *j, const double *Vre, const double *Vim)
{
    const qsim_split_desc &d = j->d;
    hipStream_t s = nullptr;
    const O4 c;
    const bool o4 = d.method == 4;
    const double dx = (d.x1 - d.x0) / d.nx;
    const double dy = d.dim == 2 ? (d.y1 - d.y0) / d.ny : 1.0;
    const cd dt(d.dt_re, d.dt_im);
    j->fused = d.boundary == 0 && fused_len_ok(d.nx, d.precision, d.dim) &&
               (d.dim == 1 || fused_len_ok(d.ny, d.precision, d.dim));
    if (j->fused && d.dim == 2) {
        j->row_var = default_variant(d.precision, d.ny, true);
        j->col_var = default_variant(d.precision, d.nx, false);
    }
    if (!j->fused && d.precision != 64)
        return set_error(QSIM_ERR_UNSUPPORTED,
                         "FP32-mixed split job needs power-of-two axes (2D: 64..8192, 1D: "
                         "64..16384) and a periodic boundary");
    // potential factors
    SplitDev &h = j->host;
    h.psi = j->psi;
    c128 *ev0, *ev1, *ev2;
    if (!o4) {
        QSIM_TRY(build_ev(j, &ev0, 0.5, Vre, Vim, s));
        h.ev[0] = ev0;
        if (j->fused) {
            QSIM_TRY(square_ev(j, &ev1, ev0, s));
            h.ev[1] = ev1;
        }
        j->ev_start = 0;
        j->ev_end = 0;
        j->nph = 1;
        j->ph_k[0] = 0;
        j->ph_e[0] = 1;
    } else {
        QSIM_TRY(build_ev(j, &ev0, c.C1, Vre, Vim, s));
        QSIM_TRY(build_ev(j, &ev1, c.C2, Vre, Vim, s));
        h.ev[0] = ev0;
        h.ev[1] = ev1;
        if (j->fused) {
            QSIM_TRY(square_ev(j, &ev2, ev0, s));
            h.ev[2] = ev2;
        }
        j->ev_start = 0;
        j->ev_end = 0;
        j->nph = 3;
        j->ph_k[0] = 0;  // D1
        j->ph_e[0] = 1;  // C2
        j->ph_k[1] = 1;  // D2
        j->ph_e[1] = 1;  // C2
        j->ph_k[2] = 0;  // D1
        j->ph_e[2] = 2;  // 2 C1 (merged across steps)
    }
    // kinetic factors
    const double Dkx = 2 * M_PI / (dx * d.nx);
    const double DTx = (Dkx * d.hbar) * (Dkx * d.hbar) / (2 * d.mass);
    const double coef[2] = {o4 ? c.D1 : 1.0, o4 ? c.D2 : 0.0};
    const int nk = o4 ? 2 : 1;
    if (d.dim == 2) {
        const double Dky = 2 * M_PI / (dy * d.ny);
        const double DTy = (Dky * d.hbar) * (Dky * d.hbar) / (2 * d.mass);
        const cd alpha = cd(-0.0, -1.0) * dt / d.hbar;  // src/SplittingMethod2D.cpp:63
        for (int k = 0; k < nk; ++k) {
            c128 *tx, *ty;
            QSIM_TRY(build_axis(j, &tx, d.nx, alpha * coef[k], DTx, 1.0 / (double)j->n));
            QSIM_TRY(build_axis(j, &ty, d.ny, alpha * coef[k], DTy, 1.0));
            h.kx[k] = tx;
            h.ky[k] = ty;
            if (!j->fused) {  // full E_T for the unfused order
                void *p;
                QSIM_TRY(dalloc(j, &p, j->n * sizeof(c128)));
                c128 *et = (c128 *)p;
                const size_t ny = d.ny, nxx = d.nx;
                const double ax = DTx, ay = DTy;
                const cd a = alpha * coef[k];
                const c128 ac = mk(a.real(), a.imag());
                QSIM_TRY(each(j->n, s, [=] __device__(size_t idx) {
                    double fi = (double)fold_half(idx / ny, nxx), fj = (double)fold_half(idx % ny, ny);
                    et[idx] = cexp(ac * (fi * fi * ax + fj * fj * ay));
                }));
                j->et_full[k] = et;
            }
        }
    } else if (d.boundary == 0) {
        // f = -I * DT * dt * tt / hbar (src/SplittingMethod1D.cpp:52-56)
        for (int k = 0; k < nk; ++k) {
            cd f = cd(-0.0, -1.0) * DTx * dt * coef[k] / d.hbar;
            c128 *t;
            QSIM_TRY(build_axis(j, &t, d.nx, f, 1.0, j->fused ? 1.0 / (double)d.nx : 1.0));
            h.kx[k] = t;
            j->et_full[k] = t;
        }
    } else {
        // wall: exp(f i^2), i < N, f = -I * DT * dt * tt / hbar, Dp = hbar pi / (dx N)
        const double Dp = d.hbar * M_PI / (dx * d.nx);
        const double DT = Dp * Dp / (2 * d.mass);
        for (int k = 0; k < nk; ++k) {
            cd f = cd(-0.0, -1.0) * DT * dt * coef[k] / d.hbar;
            std::vector<c128> hv(d.nx);
            for (size_t i = 0; i < d.nx; ++i) {
                cd e = std::exp(f * ((double)i * (double)i));
                hv[i] = mk(e.real(), e.imag());
            }
            void *p;
            QSIM_TRY(dalloc(j, &p, d.nx * sizeof(c128)));
            QSIM_HIP(hipMemcpy(p, hv.data(), d.nx * sizeof(c128), hipMemcpyHostToDevice));
            j->et_full[k] = (c128 *)p;
        }
    }
    if (j->fused) {
        if (d.precision == 32) {
            void *p;
            QSIM_TRY(dalloc(j, &p, j->n * sizeof(c64)));
            j->work = p;
        } else {
            j->work = j->psi;
        }
        h.work = j->work;
        QSIM_TRY(twiddle_table(d.nx, d.precision, &j->tw_x));
        if (d.dim == 2) QSIM_TRY(twiddle_table(d.ny, d.precision, &j->tw_y));
    } else {
        size_t dims[2] = {d.nx, d.ny};
        size_t tn = j->n;
        if (d.dim == 1 && d.boundary == 1) {
            dims[0] = 2 * d.nx;
            tn = 2 * d.nx;
        }
        QSIM_TRY(qsim_hip_fft_plan_create(&j->fwd, d.dim, dims, 0, 64));
        QSIM_TRY(qsim_hip_fft_plan_create(&j->inv, d.dim, dims, 1, 64));
        void *p;
        QSIM_TRY(dalloc(j, &p, tn * sizeof(c128)));
        j->tmp = (c128 *)p;
        QSIM_TRY(dalloc(j, &p, tn * sizeof(c128)));
        j->tmp2 = (c128 *)p;
    }
    void *p;
    QSIM_TRY(dalloc(j, &p, sizeof(SplitDev)));
    j->dev = (SplitDev *)p;
    QSIM_HIP(hipMemcpy(j->dev, &h, sizeof(SplitDev), hipMemcpyHostToDevice));
    // set-up ran on the null stream; work other streams already run is
    // not waited for (request loops create jobs while others advance)
    QSIM_HIP(hipStreamSynchronize(nullptr));
    return QSIM_OK;
}

}  // namespace

extern "C" {

int qsim_hip_split_create(qsim_split_job **job, const qsim_split_desc *d, const double *V_re,
                          const double *V_im, qsim_c128 *psi)
{
    QSIM_CHECK(job && d, "null argument");
    *job = nullptr;
    QSIM_CHECK(d->dim == 1 || d->dim == 2, "dim must be 1 or 2");
    QSIM_CHECK(d->nx >= 1 && (d->dim == 1 || d->ny >= 1), "empty grid");
    QSIM_CHECK(d->method == 2 || d->method == 4, "method must be 2 (O2) or 4 (O4)");
    QSIM_CHECK(d->boundary == 0 || (d->boundary == 1 && d->dim == 1),
               "unsupported boundary condition!");
    QSIM_CHECK(d->precision == 64 || d->precision == 32, "precision must be 64 or 32");
    QSIM_CHECK(d->mass != 0 && d->hbar != 0, "mass and hbar must be non-zero");
    QSIM_CHECK(d->x1 != d->x0 && (d->dim == 1 || d->y1 != d->y0), "empty domain");
    QSIM_CHECK(V_re && psi, "null V or psi");
    qsim_split_job *j = new qsim_split_job();
    j->d = *d;
    if (d->dim == 1) j->d.ny = 1;
    j->n = j->d.nx * j->d.ny;
    j->psi = reinterpret_cast<c128 *>(psi);
    int rc = create(j, V_re, V_im);
    if (rc != QSIM_OK) {
        destroy(j);
        return rc;
    }
    *job = j;
    return QSIM_OK;
}

int qsim_hip_split_steps(qsim_split_job *j, size_t nsteps, qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_steps");
    QSIM_CHECK(j, "null job");
    if (nsteps == 0) return QSIM_OK;
    hipStream_t s = as_stream(stream);
    if (!j->fused) return unfused_steps(j, nsteps, s);
    if (!j->use_graph || j->d.dim != 2) return fused_steps(j, j->dev, 1, nsteps, s);
    // the sweep sequence of `nsteps` steps is captured once into a HIP graph
    // and replayed (one launch instead of 2*nsteps+1); at most kMaxGraphs
    // step counts are kept per job, the oldest capture is dropped first
    auto it = j->graphs.find(nsteps);
    if (it == j->graphs.end() && j->graphs.size() >= kMaxGraphs) {
        // a replay of the graph being dropped may still be in flight
        QSIM_HIP(hipDeviceSynchronize());
        auto old = j->graphs.find(j->graph_order.front());
        (void)hipGraphExecDestroy(old->second);
        j->graphs.erase(old);
        j->graph_order.erase(j->graph_order.begin());
    }
    if (it == j->graphs.end()) {
        hipStream_t cs;
        QSIM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        int rc = QSIM_OK;
        if (e == hipSuccess) {
            rc = fused_steps(j, j->dev, 1, nsteps, cs);
            hipError_t e2 = hipStreamEndCapture(cs, &g);
            if (rc == QSIM_OK && e2 != hipSuccess) rc = hip_error(e2, "hipStreamEndCapture");
        } else {
            rc = hip_error(e, "hipStreamBeginCapture");
        }
        hipGraphExec_t ex = nullptr;
        if (rc == QSIM_OK) {
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = hip_error(e, "hipGraphInstantiate");
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        if (rc != QSIM_OK) return rc;
        it = j->graphs.emplace(nsteps, ex).first;
        j->graph_order.push_back(nsteps);
    }
    QSIM_HIP(hipGraphLaunch(it->second, s));
    return QSIM_OK;
}

int qsim_hip_split_steps_timed(qsim_split_job *j, size_t nsteps, double *kernel_ms,
                               int *kernel_launches, qsim_stream_t stream)
{
    QSIM_CHECK(j && kernel_ms && kernel_launches, "null argument");
    if (!j->fused)
        return set_error(QSIM_ERR_UNSUPPORTED, "per-kernel timing needs a fused job");
    for (int k = 0; k < TK_KINDS; ++k) {
        kernel_ms[k] = 0;
        kernel_launches[k] = 0;
    }
    if (nsteps == 0) return QSIM_OK;
    hipStream_t s = as_stream(stream);
    SweepTimer tm;
    QSIM_TRY(fused_steps(j, j->dev, 1, nsteps, s, &tm));
    QSIM_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < tm.kind.size(); ++i) {
        float ms = 0;
        QSIM_HIP(hipEventElapsedTime(&ms, tm.ev[2 * i], tm.ev[2 * i + 1]));
        kernel_ms[tm.kind[i]] += ms;
        kernel_launches[tm.kind[i]] += 1;
    }
    return QSIM_OK;
}

int qsim_hip_split_use_graph(qsim_split_job *j, int enable)
{
    QSIM_CHECK(j, "null job");
    j->use_graph = enable != 0;
    if (!j->use_graph) drop_graphs(j);
    return QSIM_OK;
}

int qsim_hip_split_set_variants(qsim_split_job *j, int row_variant, int col_variant)
{
    QSIM_CHECK(j, "null job");
    QSIM_CHECK(row_variant >= 0 && row_variant < kVariants && col_variant >= 0 &&
                   col_variant < kVariants,
               "launch variant out of range");
    j->row_var = has_variants(j->d.ny) ? row_variant : 0;
    j->col_var = has_variants(j->d.nx) ? col_variant : 0;
    drop_graphs(j);
    return QSIM_OK;
}

}  // extern "C"

namespace {

// Times the sweep variants on copies of the jobs' working states (the jobs'
// own state is untouched), with all njobs jobs in each launch as a batch runs
// them, and sets the fastest pair on every job.
int autotune_jobs(qsim_split_job *const *jobs, int njobs, int reps, int *row_variant,
                  int *col_variant, double *row_us, double *col_us, hipStream_t s)
{
    qsim_split_job *j = jobs[0];
    const int prec = j->d.precision;
    const size_t esz = prec == 64 ? sizeof(c128) : sizeof(c64);
    std::vector<void *> scratch(njobs, nullptr);
    std::vector<SplitDev> h(njobs);
    void *tab = nullptr;
    hipError_t e = hipSuccess;
    for (int i = 0; i < njobs && e == hipSuccess; ++i) {
        e = hipMalloc(&scratch[i], j->n * esz);
        if (e == hipSuccess)
            e = hipMemcpyAsync(scratch[i], jobs[i]->work, j->n * esz, hipMemcpyDeviceToDevice, s);
        h[i] = jobs[i]->host;
        h[i].work = scratch[i];
    }
    if (e == hipSuccess) e = hipMalloc(&tab, sizeof(SplitDev) * njobs);
    if (e == hipSuccess)
        e = hipMemcpy(tab, h.data(), sizeof(SplitDev) * njobs, hipMemcpyHostToDevice);
    auto release = [&]() {
        for (void *p : scratch)
            if (p) (void)hipFree(p);
        if (tab) (void)hipFree(tab);
    };
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);
        release();
        return hip_error(e, "autotune setup");
    }
    const SplitDev *jt = (const SplitDev *)tab;
    hipEvent_t a, b;
    QSIM_HIP(hipEventCreate(&a));
    QSIM_HIP(hipEventCreate(&b));
    int rc = QSIM_OK;
    // The sweeps run alternately (each reads what the other wrote, from
    // other XCDs' caches), so every candidate is timed inside alternating
    // (row, column) pairs: first the row variant with the column variant
    // fixed, then the column variant with the chosen row variant.
    int var[2] = {j->row_var, j->col_var};
    double best[2] = {1e30, 1e30};
    auto pair = [&](int vr_, int vc_) {
        int r2 =
            row_sweep(prec, j->d.ny, SW_MID, vr_, jt, njobs, j->d.nx, j->ph_e[0], j->tw_y, s);
        if (r2 == QSIM_OK)
            r2 = col_sweep(prec, j->d.nx, vc_, jt, njobs, j->d.ny, j->ph_k[0], j->tw_x, s);
        return r2;
    };
    for (int which = 0; which < 2 && rc == QSIM_OK; ++which) {
        const size_t nlen = which == 0 ? j->d.ny : j->d.nx;
        const int nv = has_variants(nlen) ? kVariants : 1;
        int chosen = var[which];
        for (int v = 0; v < nv && rc == QSIM_OK; ++v) {
            const int vr_ = which == 0 ? v : var[0], vc_ = which == 1 ? v : var[1];
            for (int w = 0; w < 2 && rc == QSIM_OK; ++w) rc = pair(vr_, vc_);
            if (rc != QSIM_OK) break;
            (void)hipEventRecord(a, s);
            for (int r = 0; r < reps && rc == QSIM_OK; ++r) rc = pair(vr_, vc_);
            (void)hipEventRecord(b, s);
            if (rc != QSIM_OK) break;
            (void)hipEventSynchronize(b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            const double us = 1e3 * ms / reps;
            if (us < best[which]) {
                best[which] = us;
                chosen = v;
            }
        }
        var[which] = chosen;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(s);
    release();
    if (rc != QSIM_OK) return rc;
    for (int i = 0; i < njobs; ++i) {
        jobs[i]->row_var = var[0];
        jobs[i]->col_var = var[1];
        drop_graphs(jobs[i]);
    }
    if (row_variant) *row_variant = var[0];
    if (col_variant) *col_variant = var[1];
    if (row_us) *row_us = best[0];
    if (col_us) *col_us = best[1];
    return QSIM_OK;
}

}  // namespace

extern "C" {

int qsim_hip_split_autotune(qsim_split_job *j, int reps, int *row_variant, int *col_variant,
                            double *row_us, double *col_us, qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_autotune");
    QSIM_CHECK(j, "null job");
    QSIM_CHECK(reps >= 1, "reps must be >= 1");
    if (!j->fused || j->d.dim != 2) {
        if (row_variant) *row_variant = 0;
        if (col_variant) *col_variant = 0;
        return QSIM_OK;
    }
    return autotune_jobs(&j, 1, reps, row_variant, col_variant, row_us, col_us,
                         as_stream(stream));
}

int qsim_hip_split_autotune_batched(qsim_split_job *const *jobs, int njobs, int reps,
                                    int *row_variant, int *col_variant, double *row_us,
                                    double *col_us, qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_autotune_batched");
    QSIM_CHECK(jobs && njobs >= 1, "null job list");
    QSIM_CHECK(reps >= 1, "reps must be >= 1");
    qsim_split_job *j0 = jobs[0];
    QSIM_CHECK(j0, "null job");
    for (int i = 1; i < njobs; ++i) {
        const qsim_split_job *ji = jobs[i];
        QSIM_CHECK(ji && ji->fused == j0->fused && ji->d.dim == j0->d.dim &&
                       ji->d.nx == j0->d.nx && ji->d.ny == j0->d.ny &&
                       ji->d.method == j0->d.method && ji->d.precision == j0->d.precision,
                   "batched autotune: jobs differ in shape, method or precision");
    }
    if (!j0->fused || j0->d.dim != 2) {
        if (row_variant) *row_variant = 0;
        if (col_variant) *col_variant = 0;
        return QSIM_OK;
    }
    return autotune_jobs(jobs, njobs, reps, row_variant, col_variant, row_us, col_us,
                         as_stream(stream));
}

int qsim_hip_split_load_psi(qsim_split_job *j, qsim_stream_t)
{
    // every qsim_hip_split_steps call starts from the caller's psi
    QSIM_CHECK(j, "null job");
    return QSIM_OK;
}

int qsim_hip_split_norm2(qsim_split_job *j, double *out, qsim_stream_t s)
{
    QSIM_CHECK(j && out, "null argument");
    return qsim_hip_norm2(reinterpret_cast<qsim_c128 *>(j->psi), j->n, out, s);
}

int qsim_hip_split_destroy(qsim_split_job *j)
{
    destroy(j);
    return QSIM_OK;
}

int qsim_hip_split_steps_batched(qsim_split_job **jobs, int njobs, size_t nsteps,
                                 qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_steps_batched");
    QSIM_CHECK(jobs && njobs >= 0, "null job list");
    if (njobs == 0 || nsteps == 0) return QSIM_OK;
    qsim_split_job *j0 = jobs[0];
    QSIM_CHECK(j0, "null job");
    bool same = j0->fused;
    for (int i = 1; i < njobs && same; ++i) {
        const qsim_split_job *ji = jobs[i];
        QSIM_CHECK(ji, "null job");
        same = ji->fused && ji->d.dim == j0->d.dim && ji->d.nx == j0->d.nx &&
               ji->d.ny == j0->d.ny && ji->d.method == j0->d.method &&
               ji->d.precision == j0->d.precision;
    }
    hipStream_t s = as_stream(stream);
    if (!same || njobs == 1) {
        for (int i = 0; i < njobs; ++i) QSIM_TRY(qsim_hip_split_steps(jobs[i], nsteps, stream));
        return QSIM_OK;
    }
    // one launch per sweep for all jobs: their descriptors in a table on the
    // current device, owned by this call (the call is synchronous; request
    // loops that reuse a table use qsim_split_batch)
    std::vector<SplitDev> h(njobs);
    for (int i = 0; i < njobs; ++i) h[i] = jobs[i]->host;
    SplitDev *table = nullptr;
    QSIM_HIP(hipMalloc(&table, sizeof(SplitDev) * njobs));
    hipError_t e = hipMemcpyAsync(table, h.data(), sizeof(SplitDev) * njobs,
                                  hipMemcpyHostToDevice, s);
    int rc = e == hipSuccess ? fused_steps(j0, table, njobs, nsteps, s)
                             : hip_error(e, "batched steps: table upload");
    // the table and the host vector must outlive the queued work
    e = hipStreamSynchronize(s);
    (void)hipFree(table);
    if (rc != QSIM_OK) return rc;
    if (e != hipSuccess) return hip_error(e, "batched steps");
    return QSIM_OK;
}

// ------------------------------------------------ reusable asynchronous batch
struct qsim_split_batch {
    int cap = 0;
    SplitDev *table = nullptr;   // device
    SplitDev *staged = nullptr;  // pinned host
    hipEvent_t copied = nullptr; // the last table upload has been read
    bool pending = false;
};

int qsim_hip_split_batch_create(qsim_split_batch **b, int capacity)
{
    QSIM_CHECK(b && capacity > 0, "batch: null out-pointer or capacity < 1");
    *b = nullptr;
    qsim_split_batch *p = new qsim_split_batch();
    p->cap = capacity;
    int rc = QSIM_OK;
    if (hipMalloc(&p->table, sizeof(SplitDev) * capacity) != hipSuccess ||
        hipHostMalloc(&p->staged, sizeof(SplitDev) * capacity, hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreateWithFlags(&p->copied, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();  // not the next launch's error
        rc = set_error(QSIM_ERR_NOMEM, "batch: allocation failed");
    }
    if (rc != QSIM_OK) {
        qsim_hip_split_batch_destroy(p);
        return rc;
    }
    *b = p;
    return QSIM_OK;
}

int qsim_hip_split_batch_steps(qsim_split_batch *b, qsim_split_job *const *jobs, int njobs,
                               size_t nsteps, qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_batch_steps");
    QSIM_CHECK(b && jobs && njobs >= 0, "batch: null argument");
    QSIM_CHECK(njobs <= b->cap, "batch: more jobs than the batch's capacity");
    if (njobs == 0 || nsteps == 0) return QSIM_OK;
    qsim_split_job *j0 = jobs[0];
    QSIM_CHECK(j0 && j0->fused, "batch: jobs must be fused split jobs");
    for (int i = 1; i < njobs; ++i) {
        const qsim_split_job *ji = jobs[i];
        QSIM_CHECK(ji && ji->fused && ji->d.dim == j0->d.dim && ji->d.nx == j0->d.nx &&
                       ji->d.ny == j0->d.ny && ji->d.method == j0->d.method &&
                       ji->d.precision == j0->d.precision && ji->row_var == j0->row_var &&
                       ji->col_var == j0->col_var,
                   "batch: jobs differ in shape, method, precision or launch variants");
    }
    hipStream_t s = as_stream(stream);
    // the pinned staging area may still feed the previous upload
    if (b->pending) QSIM_HIP(hipEventSynchronize(b->copied));
    for (int i = 0; i < njobs; ++i) b->staged[i] = jobs[i]->host;
    QSIM_HIP(hipMemcpyAsync(b->table, b->staged, sizeof(SplitDev) * njobs,
                            hipMemcpyHostToDevice, s));
    QSIM_HIP(hipEventRecord(b->copied, s));
    b->pending = true;
    return fused_steps(j0, b->table, njobs, nsteps, s);
}

int qsim_hip_split_batch_destroy(qsim_split_batch *b)
{
    if (!b) return QSIM_OK;
    if (b->pending) (void)hipEventSynchronize(b->copied);
    if (b->copied) (void)hipEventDestroy(b->copied);
    if (b->staged) (void)hipHostFree(b->staged);
    if (b->table) (void)hipFree(b->table);
    delete b;
    return QSIM_OK;
}

}  // extern "C"

// ------------------------------------------ batched steps without allocation
// qsim_hip_split_steps_batched_noalloc keeps the batch's descriptor table in
// this file-scope device array instead of a hipMalloc'd one: every GPU holds
// its own copy (kNoAllocMaxJobs * 80 B = 40 KiB) until the process ends.
// Host code never passes `scratch` itself to a kernel or a copy (in host code
// the name is a host-side stand-in, not a device address): it uses the
// address hipGetSymbolAddress returns for the current device.
namespace qsim {
constexpr int kNoAllocMaxJobs = 512;
static __device__ __align__(256) SplitDev scratch[kNoAllocMaxJobs];
}  // namespace qsim

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
SplitDev *g_noalloc_table[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int noalloc_table(SplitDev **out)
{
    int dev = 0;
    QSIM_HIP(hipGetDevice(&dev));
    QSIM_CHECK(dev >= 0 && dev < kMaxDevices, "device index beyond the library's limit");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_table[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
        g_noalloc_table[dev] = static_cast<SplitDev *>(p);
    }
    *out = g_noalloc_table[dev];
    return QSIM_OK;
}

// Descriptors travel as kernel arguments (no host staging buffer, no
// pageable copy): 32 per launch, 2.5 KiB of arguments.
constexpr int kDescPerLaunch = 32;
struct DescChunk {
    SplitDev d[kDescPerLaunch];
};

__global__ void __launch_bounds__(64) k_put_descs(SplitDev *dst, DescChunk c, int n)
{
    const int i = threadIdx.x;
    if (i < n) dst[i] = c.d[i];
}

}  // namespace

extern "C" {

int qsim_hip_split_steps_batched_noalloc(qsim_split_job *const *jobs, int njobs, size_t nsteps,
                                         qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_steps_batched_noalloc");
    QSIM_CHECK(jobs && njobs >= 0, "null job list");
    QSIM_CHECK(njobs <= kNoAllocMaxJobs,
               "qsim_hip_split_steps_batched_noalloc: more than 512 jobs in one call");
    if (njobs == 0 || nsteps == 0) return QSIM_OK;
    for (int i = 0; i < njobs; ++i) {
        QSIM_CHECK(jobs[i], "null job");
        if (!jobs[i]->fused)
            return set_error(QSIM_ERR_UNSUPPORTED,
                             "qsim_hip_split_steps_batched_noalloc: only fused split jobs "
                             "(power-of-two axes, periodic boundary)");
    }
    const qsim_split_job *j0 = jobs[0];
    bool same = true;
    for (int i = 1; i < njobs && same; ++i) {
        const qsim_split_job *ji = jobs[i];
        same = ji->d.dim == j0->d.dim && ji->d.nx == j0->d.nx && ji->d.ny == j0->d.ny &&
               ji->d.method == j0->d.method && ji->d.precision == j0->d.precision &&
               ji->row_var == j0->row_var && ji->col_var == j0->col_var;
    }
    hipStream_t s = as_stream(stream);
    if (!same || njobs == 1) {
        // each job on its own device descriptor, made when it was created
        for (int i = 0; i < njobs; ++i) QSIM_TRY(fused_steps(jobs[i], jobs[i]->dev, 1, nsteps, s));
        return QSIM_OK;
    }
    SplitDev *table = nullptr;
    QSIM_TRY(noalloc_table(&table));
    for (int b = 0; b < njobs; b += kDescPerLaunch) {
        DescChunk c;
        const int m = njobs - b < kDescPerLaunch ? njobs - b : kDescPerLaunch;
        for (int i = 0; i < m; ++i) c.d[i] = jobs[b + i]->host;
        hipLaunchKernelGGL(k_put_descs, dim3(1), dim3(64), 0, s, table + b, c, m);
        QSIM_HIP(hipGetLastError());
    }
    return fused_steps(jobs[0], table, njobs, nsteps, s);
}

}  // extern "C"
