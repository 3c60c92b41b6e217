// qsim_born.hip -- Born-series scattering job (QuPerturbation2DImpl/3DImpl on
// gfx950), many energies advanced together.
//
// Reference iteration (src/Perburbation.h:350-423, preconditioned):
//   tmp  = (Re V + i(Im V + eps)) dpsi + Re V psi0          (Add3)
//   tmp  = IFFT( FFT(tmp) / (E - T_k + i eps) ) / n          (G0)
//   dpsi = (1 - g) dpsi + g tmp,  g = slow (1 - i V/eps) ... (Vellekoop/Hao1/Hao2)
// naive (src/Perburbation.h:283-309): dpsi = IFFT( G0 FFT((dpsi + psi0) V) ) / n.
//
// Fused passes for power-of-two grids, axis order (x = axis 0 outermost):
//   inner sweep (last axis, contiguous lines): the end of iteration n
//       (IFFT_inner, /n, preconditioned update of dpsi, residual sums) and the
//       start of iteration n+1 (Add3, FFT_inner) in one read/write of
//       tmp, dpsi and V;
//   middle axis (3D only): plain FFT_y before and IFFT_y after the outer sweep;
//   outer sweep (axis 0, strided lines): FFT_x, G0 multiply, IFFT_x.
// psi0 = e^{i k.r} and the absorbing layer (src/Perburbation.h:150-177) are
// separable, so both come from per-axis tables instead of grid arrays; the
// physical V is stored once for all energies.
//
// Precision 64: complex128 iterates.  Precision 32: complex64 iterates with
// FP64 V, FP64 phase/absorber tables and FP64 residual sums.
#include <cmath>
#include <complex>
#include <vector>

#include "qsim_fft_kernels.h"
#include "qsim_fft_plan.h"
#include "qsim_reduce.h"

namespace qsim {

enum { BA_START = 0, BA_END = 1, BA_BOTH = 2 };
enum { PRE_IDENTITY = 0, PRE_VELLEKOOP = 1, PRE_HAO1 = 2, PRE_HAO2 = 3, PRE_NAIVE = 4 };

struct BornGeom {
    int dim;
    size_t n0, n1, n2;   // n2 = 1 for 2D
    size_t ninner;       // length of the last axis
    size_t nmid;         // 3D: n1; 2D: 1
};

// per-energy device descriptor
struct BornDev {
    void *dpsi;
    void *tmp;
    const c128 *pw[3];     // plane-wave factor per axis: exp(i k_a x_a)
    const double *ab[3];   // absorbing-layer Im V per axis (zeros when naive)
    double E, eps, slow, inv_n;
    double inv_eps;        // 1 / eps (0 when eps is 0)
    double a[3];           // (2 pi hbar / (d n))^2 / 2m per axis
    double *resid;         // per-block residual partial sums [2 * nblocks]
};

// potential with the absorbing layer, in the reference's summation order
__device__ __forceinline__ c128 v_eff(c128 v, const BornDev &J, int dim, size_t i, size_t j,
                                      size_t k)
{
    const double a = dim == 3 ? (ld_g(J.ab[0] + i) + ld_g(J.ab[1] + j)) + ld_g(J.ab[2] + k)
                              : ld_g(J.ab[0] + i) + ld_g(J.ab[1] + j);
    return mk(v.x, v.y + a);
}

__device__ __forceinline__ c128 psi0_at(const BornDev &J, int dim, size_t i, size_t j, size_t k)
{
    return dim == 3 ? (ld_g(J.pw[0] + i) * ld_g(J.pw[1] + j)) * ld_g(J.pw[2] + k)
                    : ld_g(J.pw[0] + i) * ld_g(J.pw[1] + j);
}

// gamma of op_Vellekoop / op_Hao1 / op_Hao2 (src/DeviceTemplate.h:1020-1077)
// with the divisions by eps done as products with 1/eps: within an ulp of the
// quotient, without two FP64 divide sequences per point in the fused sweeps.
// op_Hao2's 2 / w, w = 1 + i v / eps = (1 - v.y / eps, v.x / eps), is taken
// as 2 conj(w) / |w|^2 with one real divide instead of a complex division
// (|w|^2 >= 1 for the absorbing potentials, Im v <= 0, so the unscaled form
// neither overflows nor underflows); that also frees the registers the
// merged plane pass needs.
template <int PRE>
__device__ __forceinline__ c128 gamma_fused(c128 v, double slow, double eps, double inv_eps)
{
    if constexpr (PRE == PRE_VELLEKOOP) {
        const c128 iv = imul(v);
        return mk(1. - iv.x * inv_eps, -iv.y * inv_eps) * slow;
    } else if constexpr (PRE == PRE_HAO1) {
        const c128 iv = imul(mk(v.x, -v.y));
        return mk(1. - iv.x * inv_eps, -iv.y * inv_eps) * slow;
    } else {
        const double wr = 1. - v.y * inv_eps, wi = v.x * inv_eps;
        const double s2 = 2. * slow / (wr * wr + wi * wi);
        return mk(wr * s2, -wi * s2);
    }
}

template <int PRE>
__device__ __forceinline__ c128 update(c128 dpsi, c128 t, c128 v, double slow, double eps,
                                       double inv_eps)
{
    if constexpr (PRE == PRE_NAIVE) {
        return t;
    } else if constexpr (PRE == PRE_IDENTITY) {
        return slow == 1. ? t : dpsi * (1 - slow) + t * slow;
    } else {
        c128 g = gamma_fused<PRE>(v, slow, eps, inv_eps);
        return mk(1. - g.x, -g.y) * dpsi + g * t;
    }
}

// 1/d for normal, finite d: the v_rcp_f64 estimate, then two Newton steps
// x <- x + x (1 - d x), each of which squares the relative error, so the
// result is within an ulp of the quotient (5 instructions against the 10 of
// the IEEE divide sequence; the VALU-bound outer pass, C4 +0.5%).  Reading
// the twiddle powers from the table instead of multiplying them out was
// measured too and dropped: the loads' latency cost C4 3%.
__device__ __forceinline__ double recip_nr(double d)
{
    double x = __builtin_amdgcn_rcp(d);
    double e = fma(-d, x, 1.0);
    x = fma(x, e, x);
    e = fma(-d, x, 1.0);
    return fma(x, e, x);
}

// a * G0 with G0 = 1 / (w.x + i w.y) (op_G02D / op_G03D,
// src/DeviceTemplate.h:905-976): a * conj(w) / |w|^2 with one FP64
// reciprocal.  |w|^2 >= eps^2 > 0 and is at most a few thousand on these
// grids, so the unscaled form neither overflows nor underflows.
__device__ __forceinline__ c128 mul_g0(c128 a, double wr, double wi)
{
    const double inv = recip_nr(wr * wr + wi * wi);
    return mk((a.x * wr + a.y * wi) * inv, (a.y * wr - a.x * wi) * inv);
}

template <int PRE>
__device__ __forceinline__ c128 source(c128 dpsi, c128 p0, c128 v, double eps)
{
    if constexpr (PRE == PRE_NAIVE) return (dpsi + p0) * v;          // op_AddMul
    else return mk(v.x, v.y + eps) * dpsi + v.x * p0;                // op_Add3
}

// Launch variant of the Born sweeps per line length (LineV: 8 elements per
// thread from 128 to 4096 points).  At 128 (radices 8, 8, 2) that leaves
// registers for the prefetched operands of the inner sweep; from 256 up it
// gives one 2D solve (one energy, a few hundred lines) 4-16 times the waves
// of the default plans.  BornLFV: the strided (line-fast) passes, where
// 128-point lines keep their default plan (the 3D plane-sweep form).
template <int N>
struct BornV {
    static constexpr int V = LineV<N>::V;
};
template <int N>
struct BornLFV {
    static constexpr int V = N == 128 ? 0 : BornV<N>::V;
};

// Sum over the workgroup: wave reduction by lane shuffles, then one value
// per wave through LDS.  Every thread must call it.
template <int THREADS>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *slots)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    constexpr int NW = (THREADS + 63) / 64;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) {
        slots[w] = a;
        slots[NW + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int k = 0; k < NW; ++k) {
            sa += slots[k];
            sb += slots[NW + k];
        }
        a = sa;
        b = sb;
    }
}

// Per-energy arrays (tmp, dpsi) stream through each pass once per
// iteration, while V and the tables are read again by every energy.  With at
// least 8 MiB of them per launch, non-temporal loads keep V and the tables in
// L2 (C4: 3.01 -> 3.14e4 energy-iterations/s with the plane sweep); smaller
// launches keep normal loads, as the split sweeps do.
template <bool NT, typename T>
__device__ __forceinline__ cx<T> ld_stream(const cx<T> *p)
{
    if constexpr (NT) return ld_nt(p);
    else return ld_g(p);
}
static bool stream_nt(size_t bytes_per_launch) { return bytes_per_launch >= ((size_t)8 << 20); }

// Inner sweep: lines along the last axis.  Line g -> outer indices
// (i, j) = (g / nmid, g % nmid) in 3D, i = g in 2D.  The operands of the
// update (dpsi, V) are loaded before the inverse transform, so their latency
// hides behind it.
template <typename T, int N, int MODE, int PRE, bool NT>
__global__ void __launch_bounds__((P2Launch<T, N, false, BornV<N>::V>::THREADS))
    k_born_inner(const BornDev *__restrict__ jobs, const c128 *__restrict__ V, BornGeom geo,
                 const c128 *__restrict__ tw)
{
    using LP = P2Launch<T, N, false, BornV<N>::V>;
    using P = typename LP::Line;
    constexpr int L = LP::L;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ double slots[2 * ((LP::THREADS + 63) / 64)];
    const BornDev &J = jobs[blockIdx.y];
    const int tid = threadIdx.x;
    const int l = tid / P::TPL, t = tid % P::TPL;
    const size_t nlines = geo.n0 * geo.nmid;
    const size_t g = (size_t)blockIdx.x * L + l;
    const bool active = g < nlines;
    const size_t gg = active ? g : 0;
    const size_t base = gg * N;
    const size_t i = gg / geo.nmid, j = gg % geo.nmid;
    cx<T> *tmp = reinterpret_cast<cx<T> *>(J.tmp);
    cx<T> *dp = reinterpret_cast<cx<T> *>(J.dpsi);
    constexpr int RA = P::RA_, RZ = P::RLAST;
    const int dim = geo.dim;
    cx<T> v[P::E];
    if constexpr (MODE != BA_START) {
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                v[u * RZ + r] = active ? ld_stream<NT>(tmp + base + P::template in_pos<RZ>(t, u, r))
                                       : mk<T>(0, 0);
    }
    c128 vv[P::E];
    cx<T> dd[P::E];
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int k = P::template in_pos<RA>(t, u, r);
            vv[u * RA + r] = active ? ld_g(V + base + k) : mk(0.0, 0.0);
            dd[u * RA + r] = active ? ld_stream<NT>(dp + base + k) : mk<T>(0, 0);
        }
    if constexpr (MODE != BA_START) P::inverse(v, t, lds + l * LP::LSTR, tw);
    // per-line factors of the separable tables; the inner index varies below
    const double ab_line =
        dim == 3 ? ld_g(J.ab[0] + i) + ld_g(J.ab[1] + j) : ld_g(J.ab[0] + i);
    const c128 pw_line = dim == 3 ? ld_g(J.pw[0] + i) * ld_g(J.pw[1] + j) : ld_g(J.pw[0] + i);
    const double *ab_in = J.ab[dim == 3 ? 2 : 1];
    const c128 *pw_in = J.pw[dim == 3 ? 2 : 1];
    double rsum = 0, nsum = 0;
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int k = P::template in_pos<RA>(t, u, r);
            const c128 vk = vv[u * RA + r];
            // v_eff: (ab0 + ab1) + ab2 in 3D, ab0 + ab1 in 2D (reference order)
            const c128 ve = PRE == PRE_NAIVE ? vk : mk(vk.x, vk.y + (ab_line + ld_g(ab_in + k)));
            c128 d = cvt<double>(dd[u * RA + r]);
            if constexpr (MODE != BA_START) {
                const c128 tn = cvt<double>(v[u * RA + r]) * J.inv_n;
                const c128 nd = update<PRE>(d, tn, ve, J.slow, J.eps, J.inv_eps);
                const cx<T> nds = cvt<T>(nd);
                if (active) store_nt(&dp[base + k], nds);
                const c128 ndr = cvt<double>(nds);
                if (active) {
                    rsum += abs2(ndr - d);
                    nsum += abs2(ndr);
                }
                d = ndr;
            }
            if constexpr (MODE != BA_END) {
                const c128 p0 = pw_line * ld_g(pw_in + k);
                v[u * RA + r] = cvt<T>(source<PRE>(d, p0, ve, J.eps));
            }
        }
    if constexpr (MODE != BA_END) {
        P::forward(v, t, lds + l * LP::LSTR, tw);
        if (active) {
#pragma unroll
            for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    store_nt(&tmp[base + P::template in_pos<RZ>(t, u, r)], v[u * RZ + r]);
        }
    }
    if constexpr (MODE != BA_START) {
        // block sums of the update residual and the new norm
        block_sum2<LP::THREADS>(rsum, nsum, slots);
        if (tid == 0) {
            J.resid[2 * blockIdx.x] = rsum;
            J.resid[2 * blockIdx.x + 1] = nsum;
        }
    }
}

// y-z plane sweep (3D, FP32 iterates, n1 = n2 = 128): one workgroup per
// (x plane, energy) holds the contiguous 128 x 128 plane.  It replaces the
// middle-axis passes and the inner sweep: END runs IFFT_y, IFFT_z, the
// preconditioned update and residual sums; START runs Add3, FFT_z, FFT_y;
// BOTH runs END then START in one pass.  With a complex64 copy of V an
// iteration moves 56 B per point (BOTH 40, outer 16) instead of 96 (72 with
// END and START apart).  y lines are strided in the plane: they are loaded
// line-fast (lane -> z position), so every load and store is coalesced; the
// switch between y lines and z lines is a transpose through LDS.
constexpr int PLANE_N = 128;
using PlaneLine = Pow2Line<float, PLANE_N, 16, 8, 1, 1, true>;  // E = 16, 8 threads per line
constexpr int PLANE_THREADS = PLANE_N * PlaneLine::TPL;   // 1024
// LDS strides, each the one that took the fewest LDS cycles for its
// accesses (every store and load pattern timed under each stride mod 16 with
// rocprofv3 SQ_LDS_IDX_ACTIVE): z-line exchanges (8 lines per wave, lane =
// 8 j + t) 8 mod 16, 2.0 -> 1.0 times the conflict-free cycles against
// 1 mod 16; y-line exchanges (64 lines per wave, lane = line) odd; the
// transpose rows 8 mod 16, 1.83 -> 1.0 against PLANE_N + 1.
constexpr int PLANE_LSY = PlaneLine::LS;                  // y-line stride (1 mod 16)
constexpr int PLANE_LSZ =
    PlaneLine::LS_RAW + ((8 - PlaneLine::LS_RAW % 16) % 16 + 16) % 16;  // z lines (8 mod 16)
constexpr int PLANE_TS = PLANE_N + 8;                     // transpose row stride
constexpr int PLANE_ELEMS = cmax(PLANE_N * cmax(PLANE_LSY, PLANE_LSZ), PLANE_N * PLANE_TS);
constexpr size_t PLANE_LDS = (size_t)PLANE_ELEMS * sizeof(c64);
static_assert(PLANE_LSY % 16 == 1 && PLANE_LSZ % 16 == 8, "plane LDS strides");
static_assert(PLANE_LDS + 2 * (PLANE_THREADS / 64) * sizeof(double) <= 160 * 1024,
              "plane sweep LDS within one CU");

// The plane kernels sit at the 128-register limit of a 1024-thread
// workgroup.  They take the product-tree form of the twiddle powers: with
// the recurrence form END spilled 28 B per lane (Vellekoop, Hao1), the tree
// leaves it without spills (C4: +2.5%).  The recurrence form does leave
// BOTH at 82 registers with no spill, but C4 ran at 2.95e4
// energy-iterations/s with it against 3.58e4 with the tree's 28 B of spill.  BOTH fits the budget (Vellekoop, Hao1:
// 28 B per lane of scratch, against 220-350 B; Hao2 none) because (a) it
// reads each point's dpsi in the update loop instead of prefetching the
// line, (b) the inverse transform's outputs pass through an opaque register
// barrier before the update loop, so its last butterflies are not
// interleaved with the update's loads, and compiler barriers every two
// points keep those loads from being hoisted all at once, and (c) START
// takes its lane coordinates from an opaque copy of the thread index, so
// END's address offsets are recomputed rather than held across the pass.

template <int MODE, int PRE, bool NT>
__global__ void __launch_bounds__(PLANE_THREADS)
    k_born_plane(const BornDev *__restrict__ jobs, const c64 *__restrict__ Vf, BornGeom geo,
                 const c128 *__restrict__ tw)
{
    using P = PlaneLine;
    constexpr int N = PLANE_N;
    constexpr int RA = P::RA_, RZ = P::RLAST;
    extern __shared__ unsigned char smem_raw[];
    c64 *lds = reinterpret_cast<c64 *>(smem_raw);
    __shared__ double slots[2 * (PLANE_THREADS / 64)];
    const BornDev &J = jobs[blockIdx.y];
    const int tid = threadIdx.x;
    const size_t i = blockIdx.x;                  // x index of the plane
    const size_t pbase = i * (size_t)N * N;
    c64 *tmp = reinterpret_cast<c64 *>(J.tmp);
    c64 *dp = reinterpret_cast<c64 *>(J.dpsi);
    // z-line (row) layout: line j, lane t; y-line (line-fast) layout: z
    // position kz, lane ty
    const int j = tid / P::TPL, t = tid % P::TPL;
    const int kz = tid % N, ty = tid / N;
    // dpsi at the z-line positions the update visits, loaded before the
    // inverse transforms so its latency hides behind them (V follows later:
    // prefetching both would spill at 128 registers).  dpsi and tmp are this
    // energy's and stream through once (non-temporal loads when the launch
    // holds several energies, NT); the complex64 V is read by every energy
    // and keeps normal loads, so it stays in L2.  BOTH reads dpsi in the
    // update loop instead (a).
    constexpr bool PRE_DD = MODE != BA_BOTH;
    c64 dd[PRE_DD ? P::E : 1];
    if constexpr (PRE_DD) {
#pragma unroll
        for (int r = 0; r < RA; ++r)
            dd[r] = ld_stream<NT>(dp + pbase + (size_t)j * N + P::template in_pos<RA>(t, 0, r));
    }
    c64 v[P::E];
    if constexpr (MODE != BA_START) {
        // y lines, inverse input order
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                v[u * RZ + r] = ld_stream<NT>(tmp + pbase + (size_t)P::template in_pos<RZ>(ty, u, r) * N + kz);
        P::inverse(v, ty, lds + kz * PLANE_LSY, tw);
        // transpose: y lines (outputs at in_pos<RA>) -> z lines (inputs at in_pos<RZ>)
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RA; ++r) lds[P::template in_pos<RA>(ty, 0, r) * PLANE_TS + kz] = v[r];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                v[u * RZ + r] = lds[j * PLANE_TS + P::template in_pos<RZ>(t, u, r)];
        __syncthreads();
        P::inverse(v, t, lds + j * PLANE_LSZ, tw);
    }
    if constexpr (MODE == BA_BOTH) {  // (b)
#pragma unroll
        for (int r = 0; r < P::E; ++r) asm volatile("" : "+v"(v[r].x), "+v"(v[r].y) :: "memory");
    }
    const double ab_line = ld_g(J.ab[0] + i) + ld_g(J.ab[1] + j);
    const c128 pw_line = ld_g(J.pw[0] + i) * ld_g(J.pw[1] + j);
    double rsum = 0, nsum = 0;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
        const int k = P::template in_pos<RA>(t, 0, r);
        const c128 vk = cvt<double>(ld_g(Vf + pbase + (size_t)j * N + k));
        const c128 ve = PRE == PRE_NAIVE ? vk : mk(vk.x, vk.y + (ab_line + ld_g(J.ab[2] + k)));
        c128 d;
        if constexpr (PRE_DD) d = cvt<double>(dd[r]);
        else d = cvt<double>(ld_stream<NT>(dp + pbase + (size_t)j * N + k));
        if constexpr (MODE != BA_START) {
            const c128 tn = cvt<double>(v[r]) * J.inv_n;
            const c128 nd = update<PRE>(d, tn, ve, J.slow, J.eps, J.inv_eps);
            const c64 nds = cvt<float>(nd);
            store_nt(&dp[pbase + (size_t)j * N + k], nds);
            const c128 ndr = cvt<double>(nds);
            rsum += abs2(ndr - d);
            nsum += abs2(ndr);
            d = ndr;
        }
        if constexpr (MODE != BA_END) {
            const c128 p0 = pw_line * ld_g(J.pw[2] + k);
            v[r] = cvt<float>(source<PRE>(d, p0, ve, J.eps));
        }
        if constexpr (MODE == BA_BOTH)
            if ((r & 1) == 1) asm volatile("" ::: "memory");  // (b)
    }
    if constexpr (MODE != BA_END) {
        if constexpr (MODE != BA_START) __syncthreads();  // LDS free again
        // (c) lane coordinates from an opaque copy of the thread index
        int tid2 = tid;
        if constexpr (MODE != BA_START) asm volatile("" : "+v"(tid2));
        const int j2 = tid2 / P::TPL, t2 = tid2 % P::TPL;
        const int kz2 = tid2 % N, ty2 = tid2 / N;
        P::forward(v, t2, lds + j2 * PLANE_LSZ, tw);
        // transpose: z lines (outputs at in_pos<RZ>) -> y lines (inputs at in_pos<RA>)
        __syncthreads();
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                lds[j2 * PLANE_TS + P::template in_pos<RZ>(t2, u, r)] = v[u * RZ + r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RA; ++r) v[r] = lds[P::template in_pos<RA>(ty2, 0, r) * PLANE_TS + kz2];
        __syncthreads();
        P::forward(v, ty2, lds + kz2 * PLANE_LSY, tw);
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                store_nt(&tmp[pbase + (size_t)P::template in_pos<RZ>(ty2, u, r) * N + kz2],
                         v[u * RZ + r]);
    }
    if constexpr (MODE != BA_START) {
        block_sum2<PLANE_THREADS>(rsum, nsum, slots);
        if (tid == 0) {
            J.resid[2 * blockIdx.x] = rsum;
            J.resid[2 * blockIdx.x + 1] = nsum;
        }
    }
}

// Outer sweep: lines along axis 0, stride S = n1*n2, G0 between FFT and IFFT.
template <typename T, int N, bool NT>
__global__ void __launch_bounds__((P2Launch<T, N, true, BornLFV<N>::V>::THREADS))
    k_born_outer(const BornDev *__restrict__ jobs, BornGeom geo, const c128 *__restrict__ tw)
{
    using LP = P2Launch<T, N, true, BornLFV<N>::V>;
    using P = typename LP::Line;
    constexpr int L = LP::L;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const BornDev &J = jobs[blockIdx.y];
    const int tid = threadIdx.x;
    const int l = tid % L, t = tid / L;
    const size_t S = geo.n1 * geo.n2;
    const size_t cidx = (size_t)blockIdx.x * L + l;
    const bool active = cidx < S;
    const size_t c = active ? cidx : 0;
    cx<T> *tmp = reinterpret_cast<cx<T> *>(J.tmp);
    // fold indices of the other axes (op_G02D / op_G03D index order)
    const double fj = geo.dim == 3 ? (double)fold_half(c / geo.n2, geo.n1)
                                   : (double)fold_half(c, geo.n1);
    const double fk = geo.dim == 3 ? (double)fold_half(c % geo.n2, geo.n2) : 0.0;
    constexpr int RA = P::RA_, RZ = P::RLAST;
    cx<T> v[P::E];
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r)
            v[u * RA + r] = active ? ld_stream<NT>(tmp + c + (size_t)P::template in_pos<RA>(t, u, r) * S)
                                   : mk<T>(0, 0);
    P::forward(v, t, lds + l * LP::LSTR, tw, l);
#pragma unroll
    for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const int kk = P::template in_pos<RZ>(t, u, r);
            const double fi = (double)fold_half((size_t)kk, (size_t)N);
            const double tt = geo.dim == 3
                                  ? fi * fi * J.a[0] + fj * fj * J.a[1] + fk * fk * J.a[2]
                                  : fi * fi * J.a[0] + fj * fj * J.a[1];
            const c128 z = mul_g0(cvt<double>(v[u * RZ + r]), J.E - tt, J.eps);
            v[u * RZ + r] = cvt<T>(z);
        }
    P::inverse(v, t, lds + l * LP::LSTR, tw, l);
    if (active) {
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r)
                store_nt(&tmp[c + (size_t)P::template in_pos<RA>(t, u, r) * S], v[u * RA + r]);
    }
}

// Middle axis (3D y): plain transform of lines [n0][n1][n2] along axis 1.
template <typename T, int N, bool INV>
__global__ void __launch_bounds__((P2Launch<T, N, true, BornLFV<N>::V>::THREADS))
    k_born_mid(const BornDev *__restrict__ jobs, BornGeom geo, const c128 *__restrict__ tw)
{
    using LP = P2Launch<T, N, true, BornLFV<N>::V>;
    using P = typename LP::Line;
    constexpr int L = LP::L;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const BornDev &J = jobs[blockIdx.y];
    const int tid = threadIdx.x;
    const int l = tid % L, t = tid / L;
    const size_t S = geo.n2;
    const size_t nlines = geo.n0 * geo.n2;
    const size_t g = (size_t)blockIdx.x * L + l;
    const bool active = g < nlines;
    const size_t base = active ? line_base(g, N, S) : 0;
    cx<T> *tmp = reinterpret_cast<cx<T> *>(J.tmp);
    constexpr int R0 = INV ? P::RLAST : P::RA_;
    constexpr int R1 = INV ? P::RA_ : P::RLAST;
    cx<T> v[P::E];
#pragma unroll
    for (int u = 0; u < P::E / R0; ++u)
#pragma unroll
        for (int r = 0; r < R0; ++r)
            v[u * R0 + r] = active ? ld_g(tmp + base + (size_t)P::template in_pos<R0>(t, u, r) * S)
                                   : mk<T>(0, 0);
    if constexpr (INV)
        P::inverse(v, t, lds + l * LP::LSTR, tw, l);
    else
        P::forward(v, t, lds + l * LP::LSTR, tw, l);
    if (active) {
#pragma unroll
        for (int u = 0; u < P::E / R1; ++u)
#pragma unroll
            for (int r = 0; r < R1; ++r)
                store_nt(&tmp[base + (size_t)P::template in_pos<R1>(t, u, r) * S], v[u * R1 + r]);
    }
}

template <typename F>
__global__ void __launch_bounds__(256) k_each_b(size_t n, F f)
{
    size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) f(i);
}

template <typename F>
int each_b(size_t n, hipStream_t s, F f)
{
    if (!n) return QSIM_OK;
    size_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_each_b<F>, dim3((unsigned)b), dim3(256), 0, s, n, f);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

using cd = std::complex<double>;

}  // namespace qsim

using namespace qsim;

struct qsim_born_job {
    qsim_born_desc d{};
    BornGeom geo{};
    size_t n = 0;
    int nE = 0;
    int pre = PRE_NAIVE;
    bool fused = false;
    int prec = 64;
    const c128 *V = nullptr;    // caller's physical potential (device)
    std::vector<double> E;
    std::vector<BornDev> host;  // per energy
    BornDev *dev = nullptr;     // device copy of `host`
    BornDev *active_tab = nullptr;
    BornDev *tab_host = nullptr;  // pinned staging of the active table
    double *rtot_dev = nullptr;   // per active energy: residual and norm totals
    double *rtot_host = nullptr;  // pinned copy of rtot_dev
    std::vector<void *> owned;
    std::vector<bool> started;  // tmp holds a started iteration
    std::vector<int> iters;
    std::vector<double> resid;
    size_t inner_blocks = 0;
    // 3D FP32 jobs with 128 x 128 y-z planes run the plane sweep; it reads a
    // complex64 copy of V
    bool plane = false;
    c64 *Vf = nullptr;
    const void *tw[3] = {nullptr, nullptr, nullptr};
    // unfused path (FP64, any size)
    qsim_fft_plan *fwd = nullptr, *inv = nullptr;
    c128 *psi0 = nullptr, *veff = nullptr, *last = nullptr;
};

namespace {

int balloc(qsim_born_job *j, void **p, size_t bytes)
{
    QSIM_HIP(hipMalloc(p, bytes));
    j->owned.push_back(*p);
    return QSIM_OK;
}

template <typename T>
int upload(qsim_born_job *j, const std::vector<T> &h, const T **out)
{
    void *p;
    QSIM_TRY(balloc(j, &p, h.size() * sizeof(T)));
    QSIM_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T *)p;
    return QSIM_OK;
}

// cal_v15 (src/Perburbation.h:45-51)
double cal_v15(double x, double k0, double e0)
{
    x *= k0;
    return (-0.8 * e0) *
           ((1 - x * x / (15. * 15)) > 0 ? std::exp(-0.015 * x * x * 1. / std::sqrt(1 - x * x / (15. * 15)))
                                         : 0);
}

size_t pfold(size_t i, size_t n) { return (i < n / 2) ? i : n - i; }

bool born_len_ok(size_t n, int prec)
{
    if (n < 64 || (n & (n - 1))) return false;
    return prec == 64 ? n <= 4096 : n <= 8192;
}

#define QSIM_BSWITCH(PREC, NVAL, CALL)                                                \
    switch (NVAL) {                                                                   \
        case 64: return PREC == 64 ? CALL(double, 64) : CALL(float, 64);              \
        case 128: return PREC == 64 ? CALL(double, 128) : CALL(float, 128);           \
        case 256: return PREC == 64 ? CALL(double, 256) : CALL(float, 256);           \
        case 512: return PREC == 64 ? CALL(double, 512) : CALL(float, 512);           \
        case 1024: return PREC == 64 ? CALL(double, 1024) : CALL(float, 1024);        \
        case 2048: return PREC == 64 ? CALL(double, 2048) : CALL(float, 2048);        \
        case 4096: return PREC == 64 ? CALL(double, 4096) : CALL(float, 4096);        \
        default: break;                                                               \
    }                                                                                 \
    return set_error(QSIM_ERR_UNSUPPORTED, "no fused Born kernel for this length");

template <typename T, int N, int MODE, int PRE>
int launch_inner_t(const BornDev *tab, int nact, const c128 *V, const BornGeom &geo,
                   const void *tw, hipStream_t s)
{
    using LP = P2Launch<T, N, false, BornV<N>::V>;
    size_t nlines = geo.n0 * geo.nmid;
    // V is what the hint protects, and only other energies of the same
    // launch re-read it soon: complex128 128^3 x 8 energies 1.54 -> 1.61e4
    // energy-iterations/s, one energy unchanged (-1%) so it keeps plain loads
    const bool nt = nact >= 2 && stream_nt((size_t)nact * nlines * N * sizeof(cx<T>));
    auto fn = nt ? k_born_inner<T, N, MODE, PRE, true> : k_born_inner<T, N, MODE, PRE, false>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    dim3 grid((unsigned)((nlines + LP::L - 1) / LP::L), (unsigned)nact);
    hipLaunchKernelGGL(fn, grid, dim3(LP::THREADS), LP::LDS_BYTES, s, tab, V, geo,
                       (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

template <typename T, int N>
size_t inner_blocks_t(const BornGeom &geo)
{
    using LP = P2Launch<T, N, false, BornV<N>::V>;
    return (geo.n0 * geo.nmid + LP::L - 1) / LP::L;
}

template <typename T, int N, int MODE>
int launch_inner_m(int pre, const BornDev *tab, int nact, const c128 *V, const BornGeom &geo,
                   const void *tw, hipStream_t s)
{
    switch (pre) {
        case PRE_IDENTITY: return launch_inner_t<T, N, MODE, PRE_IDENTITY>(tab, nact, V, geo, tw, s);
        case PRE_VELLEKOOP: return launch_inner_t<T, N, MODE, PRE_VELLEKOOP>(tab, nact, V, geo, tw, s);
        case PRE_HAO1: return launch_inner_t<T, N, MODE, PRE_HAO1>(tab, nact, V, geo, tw, s);
        case PRE_HAO2: return launch_inner_t<T, N, MODE, PRE_HAO2>(tab, nact, V, geo, tw, s);
        default: return launch_inner_t<T, N, MODE, PRE_NAIVE>(tab, nact, V, geo, tw, s);
    }
}

template <typename T, int N>
int launch_inner_n(int mode, int pre, const BornDev *tab, int nact, const c128 *V,
                   const BornGeom &geo, const void *tw, hipStream_t s)
{
    if (mode == BA_START) return launch_inner_m<T, N, BA_START>(pre, tab, nact, V, geo, tw, s);
    if (mode == BA_END) return launch_inner_m<T, N, BA_END>(pre, tab, nact, V, geo, tw, s);
    return launch_inner_m<T, N, BA_BOTH>(pre, tab, nact, V, geo, tw, s);
}

template <int MODE, int PRE>
int launch_plane_t(const BornDev *tab, int nact, const c64 *Vf, const BornGeom &geo,
                   const void *tw, hipStream_t s)
{
    // non-temporal per-energy loads when other energies of the launch re-read V
    auto fn = nact >= 2 ? k_born_plane<MODE, PRE, true> : k_born_plane<MODE, PRE, false>;
    QSIM_TRY(lds_limit((const void *)fn, PLANE_LDS));
    dim3 grid((unsigned)geo.n0, (unsigned)nact);
    hipLaunchKernelGGL(fn, grid, dim3(PLANE_THREADS), PLANE_LDS, s, tab, Vf, geo,
                       (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

template <int MODE>
int launch_plane_m(int pre, const BornDev *tab, int nact, const c64 *Vf, const BornGeom &geo,
                   const void *tw, hipStream_t s)
{
    switch (pre) {
        case PRE_IDENTITY: return launch_plane_t<MODE, PRE_IDENTITY>(tab, nact, Vf, geo, tw, s);
        case PRE_VELLEKOOP: return launch_plane_t<MODE, PRE_VELLEKOOP>(tab, nact, Vf, geo, tw, s);
        case PRE_HAO1: return launch_plane_t<MODE, PRE_HAO1>(tab, nact, Vf, geo, tw, s);
        case PRE_HAO2: return launch_plane_t<MODE, PRE_HAO2>(tab, nact, Vf, geo, tw, s);
        default: return launch_plane_t<MODE, PRE_NAIVE>(tab, nact, Vf, geo, tw, s);
    }
}

int plane(int mode, int pre, const BornDev *tab, int nact, const c64 *Vf, const BornGeom &geo,
          const void *tw, hipStream_t s)
{
    if (geo.dim != 3 || geo.n1 != (size_t)PLANE_N || geo.n2 != (size_t)PLANE_N)
        return set_error(QSIM_ERR_UNSUPPORTED, "plane sweep needs 128 x 128 y-z planes");
    if (mode == BA_START) return launch_plane_m<BA_START>(pre, tab, nact, Vf, geo, tw, s);
    // end + start in one pass (40 B per point instead of 56)
    if (mode == BA_BOTH) return launch_plane_m<BA_BOTH>(pre, tab, nact, Vf, geo, tw, s);
    QSIM_TRY(launch_plane_m<BA_END>(pre, tab, nact, Vf, geo, tw, s));
    if (mode == BA_END) return QSIM_OK;
    return launch_plane_m<BA_START>(pre, tab, nact, Vf, geo, tw, s);
}

int inner(int prec, int mode, int pre, const BornDev *tab, int nact, const c128 *V,
          const BornGeom &geo, const void *tw, hipStream_t s)
{
#define INNERCALL(T, NN) launch_inner_n<T, NN>(mode, pre, tab, nact, V, geo, tw, s)
    QSIM_BSWITCH(prec, geo.ninner, INNERCALL)
#undef INNERCALL
}

size_t inner_blocks(int prec, const BornGeom &geo)
{
#define IB(NN) \
    case NN: return prec == 64 ? inner_blocks_t<double, NN>(geo) : inner_blocks_t<float, NN>(geo);
    switch (geo.ninner) {
        IB(64) IB(128) IB(256) IB(512) IB(1024) IB(2048) IB(4096)
        default: return 0;
    }
#undef IB
}

template <typename T, int N>
int launch_outer_t(const BornDev *tab, int nact, const BornGeom &geo, const void *tw,
                   hipStream_t s)
{
    using LP = P2Launch<T, N, true, BornLFV<N>::V>;
    size_t S = geo.n1 * geo.n2;
    auto fn = stream_nt((size_t)nact * S * N * sizeof(cx<T>)) ? k_born_outer<T, N, true>
                                                               : k_born_outer<T, N, false>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    dim3 grid((unsigned)((S + LP::L - 1) / LP::L), (unsigned)nact);
    hipLaunchKernelGGL(fn, grid, dim3(LP::THREADS), LP::LDS_BYTES, s, tab, geo, (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

int outer(int prec, const BornDev *tab, int nact, const BornGeom &geo, const void *tw,
          hipStream_t s)
{
#define OUTERCALL(T, NN) launch_outer_t<T, NN>(tab, nact, geo, tw, s)
    QSIM_BSWITCH(prec, geo.n0, OUTERCALL)
#undef OUTERCALL
}

template <typename T, int N, bool INV>
int launch_mid_t(const BornDev *tab, int nact, const BornGeom &geo, const void *tw,
                 hipStream_t s)
{
    using LP = P2Launch<T, N, true, BornLFV<N>::V>;
    auto fn = k_born_mid<T, N, INV>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    size_t nlines = geo.n0 * geo.n2;
    dim3 grid((unsigned)((nlines + LP::L - 1) / LP::L), (unsigned)nact);
    hipLaunchKernelGGL(fn, grid, dim3(LP::THREADS), LP::LDS_BYTES, s, tab, geo, (const c128 *)tw);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

int mid(int prec, bool inv, const BornDev *tab, int nact, const BornGeom &geo, const void *tw,
        hipStream_t s)
{
#define MIDCALL(T, NN) (inv ? launch_mid_t<T, NN, true>(tab, nact, geo, tw, s) \
                            : launch_mid_t<T, NN, false>(tab, nact, geo, tw, s))
    QSIM_BSWITCH(prec, geo.n1, MIDCALL)
#undef MIDCALL
}

void destroy(qsim_born_job *j)
{
    if (!j) return;
    QSIM_DEVICE_WAIT_GUARD;
    for (void *p : j->owned) (void)hipFree(p);
    if (j->tab_host) (void)hipHostFree(j->tab_host);
    if (j->rtot_host) (void)hipHostFree(j->rtot_host);
    if (j->fwd) qsim_hip_fft_plan_destroy(j->fwd);
    if (j->inv) qsim_hip_fft_plan_destroy(j->inv);
    delete j;
}

// epsilon rule (src/Perburbation.h:329-344) over V_eff of energy e, reduced
// on the fly (V_eff = V + i absorbing layer from the per-axis tables)
int epsilon_for(qsim_born_job *j, int e, double *eps)
{
    const BornDev J = j->host[e];
    const BornGeom geo = j->geo;
    const c128 *V = j->V;
    const size_t n = j->n;
    auto ve_at = [=] __device__(size_t idx) {
        size_t i, jj, k;
        if (geo.dim == 3) {
            k = idx % geo.n2;
            jj = (idx / geo.n2) % geo.n1;
            i = idx / (geo.n1 * geo.n2);
        } else {
            k = 0;
            jj = idx % geo.n1;
            i = idx / geo.n1;
        }
        return v_eff(V[idx], J, geo.dim, i, jj, k);
    };
    double minEps = 0;
    if (j->pre == PRE_IDENTITY) {
        double r[2];
        QSIM_TRY(red::reduce_sum<2>(n, nullptr, r, [=] __device__(size_t idx, double *v) {
            const c128 ve = ve_at(idx);
            v[0] = ve.y;
            v[1] = abs2(ve);
        }));
        minEps = 0.5 * r[1] / (-r[0]);
    } else {
        double m2 = 0;
        QSIM_TRY(red::reduce_ext<true>(n, nullptr, 0.0, &m2,
                                       [=] __device__(size_t idx) { return abs2(ve_at(idx)); }));
        minEps = std::sqrt(m2);
    }
    if (*eps < minEps) *eps = minEps;
    return QSIM_OK;
}

int create(qsim_born_job *j)
{
    const qsim_born_desc &d = j->d;
    const int dim = d.dim;
    BornGeom &g = j->geo;
    g.dim = dim;
    g.n0 = d.n[0];
    g.n1 = d.n[1];
    g.n2 = dim == 3 ? d.n[2] : 1;
    g.ninner = dim == 3 ? g.n2 : g.n1;
    g.nmid = dim == 3 ? g.n1 : 1;
    j->n = g.n0 * g.n1 * g.n2;
    j->pre = d.preconditional ? d.preconditioner : PRE_NAIVE;
    j->prec = d.precision;
    j->fused = born_len_ok(g.n0, j->prec) && born_len_ok(g.ninner, j->prec) &&
               (dim == 2 || born_len_ok(g.n1, j->prec));
    if (!j->fused && j->prec != 64)
        return set_error(QSIM_ERR_UNSUPPORTED,
                         "FP32 Born job needs power-of-two axes (64..4096)");
    double dx[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) dx[a] = (d.hi[a] - d.lo[a]) / d.n[a];
    double nm2 = 0;
    for (int a = 0; a < dim; ++a) nm2 += d.direction[a] * d.direction[a];
    QSIM_CHECK(nm2 > 0, "zero incident direction");
    const double nm = 1 / std::sqrt(nm2);
    const size_t esz = j->prec == 64 ? sizeof(c128) : sizeof(c64);
    j->host.resize(j->nE);
    j->started.assign(j->nE, false);
    j->iters.assign(j->nE, 0);
    j->resid.assign(j->nE, 0.0);
    j->plane = j->fused && j->prec == 32 && dim == 3 && g.n1 == (size_t)PLANE_N &&
               g.n2 == (size_t)PLANE_N;
    j->inner_blocks = j->plane ? g.n0 : (j->fused ? inner_blocks(j->prec, g) : 0);
    for (int e = 0; e < j->nE; ++e) {
        BornDev &J = j->host[e];
        const double En = j->E[e];
        QSIM_CHECK(En > 0, "energy must be positive");
        // k0 = sqrt(2 m E) / hbar, src/Perturbation3DImpl.cpp:14 + ScatteringSolverImpl.cpp:10
        const double k0 = std::sqrt(2 * En * d.mass) / d.hbar;
        if (d.preconditional) {
            if (dim == 3) {
                const double lambda = 20 / (std::sqrt(2 * d.mass * En) / d.hbar);
                for (int a = 0; a < 3; ++a)
                    if (d.n[a] * dx[a] < lambda)
                        return set_error(QSIM_ERR_INVALID,
                                         "too small size: lambda: " + std::to_string(lambda));
            } else {
                const double lambda = 2 * M_PI / (std::sqrt(2 * d.mass * En) / d.hbar);
                for (int a = 0; a < 2; ++a)
                    if (d.n[a] * dx[a] < 24 * lambda)
                        return set_error(QSIM_ERR_INVALID, "too small size");
            }
        }
        for (int a = 0; a < dim; ++a) {
            const double ka = k0 * (d.direction[a] * nm);
            std::vector<c128> pw(d.n[a]);
            std::vector<double> ab(d.n[a], 0.0);
            for (size_t i = 0; i < d.n[a]; ++i) {
                const double x = d.lo[a] + dx[a] * i;
                const double ph = ka * x;
                pw[i] = mk(std::cos(ph), std::sin(ph));
                if (d.preconditional) ab[i] = cal_v15(pfold(i, d.n[a]) * dx[a], k0, En);
            }
            QSIM_TRY(upload(j, pw, &J.pw[a]));
            QSIM_TRY(upload(j, ab, &J.ab[a]));
            // CalDp2Div2M, src/Perburbation.h:16-22
            const double Dp = 2 * M_PI / (dx[a] * d.n[a]) * d.hbar;
            J.a[a] = Dp * Dp / (2 * d.mass);
        }
        if (dim == 2) {
            J.pw[2] = J.pw[1];
            J.ab[2] = J.ab[1];
            J.a[2] = 0;
        }
        J.E = En;
        J.slow = d.slow;
        J.inv_n = 1. / (double)j->n;
        double eps = d.epsilon;
        if (d.preconditional) QSIM_TRY(epsilon_for(j, e, &eps));
        J.eps = eps;
        J.inv_eps = eps != 0 ? 1.0 / eps : 0.0;
        void *p;
        QSIM_TRY(balloc(j, &p, j->n * esz));
        J.dpsi = p;
        QSIM_HIP(hipMemset(p, 0, j->n * esz));
        QSIM_TRY(balloc(j, &p, j->n * esz));
        J.tmp = p;
        size_t nb = j->fused ? j->inner_blocks : 1;
        QSIM_TRY(balloc(j, &p, 2 * nb * sizeof(double)));
        J.resid = (double *)p;
    }
    void *p;
    if (j->plane) {
        QSIM_TRY(balloc(j, &p, j->n * sizeof(c64)));
        j->Vf = (c64 *)p;
        c64 *vf = j->Vf;
        const c128 *V = j->V;
        QSIM_TRY(each_b(j->n, nullptr, [=] __device__(size_t idx) { vf[idx] = cvt<float>(V[idx]); }));
    }
    QSIM_TRY(balloc(j, &p, sizeof(BornDev) * j->nE));
    j->dev = (BornDev *)p;
    QSIM_HIP(hipMemcpy(j->dev, j->host.data(), sizeof(BornDev) * j->nE, hipMemcpyHostToDevice));
    QSIM_TRY(balloc(j, &p, sizeof(BornDev) * j->nE));
    j->active_tab = (BornDev *)p;
    QSIM_HIP(hipHostMalloc(&p, sizeof(BornDev) * j->nE, hipHostMallocDefault));
    j->tab_host = (BornDev *)p;
    QSIM_TRY(balloc(j, &p, 2 * sizeof(double) * j->nE));
    j->rtot_dev = (double *)p;
    QSIM_HIP(hipHostMalloc(&p, 2 * sizeof(double) * j->nE, hipHostMallocDefault));
    j->rtot_host = (double *)p;
    if (j->fused) {
        QSIM_TRY(twiddle_table(g.n0, j->prec, &j->tw[0]));
        QSIM_TRY(twiddle_table(g.n1, j->prec, &j->tw[1]));
        if (dim == 3) QSIM_TRY(twiddle_table(g.n2, j->prec, &j->tw[2]));
    } else {
        size_t dims[3] = {d.n[0], d.n[1], d.n[2]};
        QSIM_TRY(qsim_hip_fft_plan_create(&j->fwd, dim, dims, 0, 64));
        QSIM_TRY(qsim_hip_fft_plan_create(&j->inv, dim, dims, 1, 64));
        QSIM_TRY(balloc(j, &p, j->n * sizeof(c128)));
        j->psi0 = (c128 *)p;
        QSIM_TRY(balloc(j, &p, j->n * sizeof(c128)));
        j->veff = (c128 *)p;
        QSIM_TRY(balloc(j, &p, j->n * sizeof(c128)));
        j->last = (c128 *)p;
    }
    // set-up ran on the null stream; work other streams already run is
    // not waited for (request loops create jobs while others advance)
    QSIM_HIP(hipStreamSynchronize(nullptr));
    return QSIM_OK;
}

int fused_pass_set(qsim_born_job *j, int mode, const BornDev *tab, int nact, hipStream_t s)
{
    // mode: BA_START or BA_BOTH starts an iteration and runs it up to the
    // point where only the inner inverse + update remains
    const void *tw_in = j->geo.dim == 3 ? j->tw[2] : j->tw[1];
    if (j->plane) {
        QSIM_TRY(plane(mode, j->pre, tab, nact, j->Vf, j->geo, tw_in, s));
        return outer(j->prec, tab, nact, j->geo, j->tw[0], s);
    }
    QSIM_TRY(inner(j->prec, mode, j->pre, tab, nact, j->V, j->geo, tw_in, s));
    if (j->geo.dim == 3) QSIM_TRY(mid(j->prec, false, tab, nact, j->geo, j->tw[1], s));
    QSIM_TRY(outer(j->prec, tab, nact, j->geo, j->tw[0], s));
    if (j->geo.dim == 3) QSIM_TRY(mid(j->prec, true, tab, nact, j->geo, j->tw[1], s));
    return QSIM_OK;
}

// Totals of the per-block residual sums of each active energy (tab order),
// one thread per energy adding the blocks in block order.
__global__ void k_resid_total(const BornDev *__restrict__ tab, int nact, size_t nb,
                              double *__restrict__ out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nact) return;
    const double *p = tab[a].resid;
    double r = 0, nn = 0;
    for (size_t b = 0; b < nb; ++b) {
        r += p[2 * b];
        nn += p[2 * b + 1];
    }
    out[2 * a] = r;
    out[2 * a + 1] = nn;
}

// Relative update of each active energy, sqrt(sum |d new - d old|^2 /
// sum |d new|^2); `act` lists the energies of the device's active table in
// its order.  One small kernel and one copy, then a wait on the stream.
int read_residuals(qsim_born_job *j, const std::vector<int> &act, hipStream_t s,
                   std::vector<double> &out)
{
    const int nact = (int)act.size();
    out.assign(act.size(), 0.0);
    if (nact == 0) return QSIM_OK;
    hipLaunchKernelGGL(k_resid_total, dim3((nact + 63) / 64), dim3(64), 0, s,
                       (const BornDev *)j->active_tab, nact, j->inner_blocks, j->rtot_dev);
    QSIM_HIP(hipGetLastError());
    QSIM_HIP(hipMemcpyAsync(j->rtot_host, j->rtot_dev, 2 * sizeof(double) * nact,
                            hipMemcpyDeviceToHost, s));
    QSIM_HIP(hipStreamSynchronize(s));
    for (int a = 0; a < nact; ++a) {
        const double r = j->rtot_host[2 * a], nn = j->rtot_host[2 * a + 1];
        out[a] = nn > 0 ? std::sqrt(r / nn) : 0.0;
    }
    return QSIM_OK;
}

int fused_iterate(qsim_born_job *j, int max_iter, double tol, hipStream_t s)
{
    std::vector<int> act;
    for (int e = 0; e < j->nE; ++e) act.push_back(e);
    std::vector<int> done_it(j->nE, 0);
    const int check_every = tol > 0 ? 4 : 0;
    int it = 0;
    auto upload_tab = [&]() -> int {
        // the previous upload has completed (each one ends with a stream sync)
        for (size_t a = 0; a < act.size(); ++a) j->tab_host[a] = j->host[act[a]];
        QSIM_HIP(hipMemcpyAsync(j->active_tab, j->tab_host, sizeof(BornDev) * act.size(),
                                hipMemcpyHostToDevice, s));
        QSIM_HIP(hipStreamSynchronize(s));
        return QSIM_OK;
    };
    QSIM_TRY(upload_tab());
    // every energy of one job starts and stops on the same iteration count
    // unless a tolerance removes it from the active table
    bool started = j->started[0];
    while (it < max_iter && !act.empty()) {
        const int nact = (int)act.size();
        QSIM_TRY(fused_pass_set(j, started ? BA_BOTH : BA_START, j->active_tab, nact, s));
        started = true;
        ++it;
        for (int e : act) {
            j->iters[e] += 1;
            done_it[e] += 1;
        }
        if (check_every && it > 1 && (it % check_every) == 0) {
            // the residuals written by the BOTH pass belong to the previous update
            std::vector<double> r;
            QSIM_TRY(read_residuals(j, act, s, r));
            std::vector<int> keep;
            for (size_t a = 0; a < act.size(); ++a) {
                j->resid[act[a]] = r[a];
                if (r[a] >= tol) keep.push_back(act[a]);
                else {
                    // converged: its started iteration is discarded (dpsi is final)
                    j->started[act[a]] = false;
                    j->iters[act[a]] -= 1;
                }
            }
            if (keep.size() != act.size()) {
                act = keep;
                if (!act.empty()) QSIM_TRY(upload_tab());
            }
        }
    }
    if (!act.empty()) {
        // finish the started iteration of every still-active energy
        if (j->plane)
            QSIM_TRY(plane(BA_END, j->pre, j->active_tab, (int)act.size(), j->Vf, j->geo,
                           j->tw[2], s));
        else
            QSIM_TRY(inner(j->prec, BA_END, j->pre, j->active_tab, (int)act.size(), j->V,
                           j->geo, j->geo.dim == 3 ? j->tw[2] : j->tw[1], s));
        std::vector<double> r;
        QSIM_TRY(read_residuals(j, act, s, r));
        for (size_t a = 0; a < act.size(); ++a) {
            j->resid[act[a]] = r[a];
            j->started[act[a]] = false;
        }
    }
    for (int e = 0; e < j->nE; ++e) j->started[e] = false;
    return QSIM_OK;
}

// Unfused FP64 reference-order iteration for any grid the FFT plans accept.
int unfused_iterate(qsim_born_job *j, int max_iter, double tol, hipStream_t s)
{
    const size_t n = j->n;
    const BornGeom geo = j->geo;
    const c128 *V = j->V;
    for (int e = 0; e < j->nE; ++e) {
        const BornDev J = j->host[e];
        c128 *dp = (c128 *)J.dpsi, *tmp = (c128 *)J.tmp, *p0 = j->psi0, *ve = j->veff;
        const int pre = j->pre;
        QSIM_TRY(each_b(n, s, [=] __device__(size_t idx) {
            size_t i, jj, k;
            if (geo.dim == 3) {
                k = idx % geo.n2;
                jj = (idx / geo.n2) % geo.n1;
                i = idx / (geo.n1 * geo.n2);
            } else {
                k = 0;
                jj = idx % geo.n1;
                i = idx / geo.n1;
            }
            p0[idx] = psi0_at(J, geo.dim, i, jj, k);
            ve[idx] = pre == PRE_NAIVE ? V[idx] : v_eff(V[idx], J, geo.dim, i, jj, k);
        }));
        for (int it = 0; it < max_iter; ++it) {
            QSIM_HIP(hipMemcpyAsync(j->last, dp, n * sizeof(c128), hipMemcpyDeviceToDevice, s));
            const double eps = J.eps;
            if (pre == PRE_NAIVE)
                QSIM_TRY(each_b(n, s, [=] __device__(size_t i) { tmp[i] = (dp[i] + p0[i]) * ve[i]; }));
            else
                QSIM_TRY(each_b(n, s, [=] __device__(size_t i) {
                    c128 v = ve[i];
                    tmp[i] = mk(v.x, v.y + eps) * dp[i] + v.x * p0[i];
                }));
            QSIM_TRY(qsim_hip_fft_exec(j->fwd, tmp, tmp, s));
            if (geo.dim == 3)
                QSIM_TRY(qsim_hip_g03d((qsim_c128 *)tmp, J.E, eps, J.a[0], J.a[1], J.a[2], geo.n0,
                                       geo.n1, geo.n2, s));
            else
                QSIM_TRY(qsim_hip_g02d((qsim_c128 *)tmp, J.E, eps, J.a[0], J.a[1], geo.n0, geo.n1, s));
            QSIM_TRY(qsim_hip_fft_exec(j->inv, tmp, tmp, s));
            const double in = J.inv_n, slow = J.slow, ie = J.inv_eps;
            switch (pre) {
                case PRE_NAIVE:
                    QSIM_TRY(each_b(n, s, [=] __device__(size_t i) { dp[i] = update<PRE_NAIVE>(dp[i], tmp[i] * in, ve[i], slow, eps, ie); }));
                    break;
                case PRE_IDENTITY:
                    QSIM_TRY(each_b(n, s, [=] __device__(size_t i) { dp[i] = update<PRE_IDENTITY>(dp[i], tmp[i] * in, ve[i], slow, eps, ie); }));
                    break;
                case PRE_VELLEKOOP:
                    QSIM_TRY(each_b(n, s, [=] __device__(size_t i) { dp[i] = update<PRE_VELLEKOOP>(dp[i], tmp[i] * in, ve[i], slow, eps, ie); }));
                    break;
                case PRE_HAO1:
                    QSIM_TRY(each_b(n, s, [=] __device__(size_t i) { dp[i] = update<PRE_HAO1>(dp[i], tmp[i] * in, ve[i], slow, eps, ie); }));
                    break;
                default:
                    QSIM_TRY(each_b(n, s, [=] __device__(size_t i) { dp[i] = update<PRE_HAO2>(dp[i], tmp[i] * in, ve[i], slow, eps, ie); }));
                    break;
            }
            j->iters[e] += 1;
            double num = 0, den = 0;
            QSIM_TRY(qsim_hip_minus_norm2((qsim_c128 *)dp, (qsim_c128 *)j->last, n, &num, s));
            QSIM_TRY(qsim_hip_norm2((qsim_c128 *)dp, n, &den, s));
            j->resid[e] = den > 0 ? std::sqrt(num / den) : 0.0;
            if (tol > 0 && j->resid[e] < tol) break;
        }
    }
    return QSIM_OK;
}

}  // namespace

extern "C" {

int qsim_hip_born_create(qsim_born_job **job, const qsim_born_desc *d, const qsim_c128 *V,
                         const double *energies, int nE)
{
    QSIM_CHECK(job && d && V && energies, "null argument");
    *job = nullptr;
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(d->dim == 2 || d->dim == 3, "dim must be 2 or 3");
    QSIM_CHECK(nE >= 1, "need at least one energy");
    QSIM_CHECK(d->precision == 64 || d->precision == 32, "precision must be 64 or 32");
    QSIM_CHECK(!d->preconditional || (d->preconditioner >= 0 && d->preconditioner <= 3),
               "Unknown preconditioner");
    QSIM_CHECK(d->mass != 0 && d->hbar != 0, "mass and hbar must be non-zero");
    for (int a = 0; a < d->dim; ++a) {
        QSIM_CHECK(d->n[a] >= 1, "empty grid");
        QSIM_CHECK(d->hi[a] != d->lo[a], "empty domain");
    }
    qsim_born_job *j = new qsim_born_job();
    j->d = *d;
    if (d->dim == 2) j->d.n[2] = 1;
    j->V = reinterpret_cast<const c128 *>(V);
    j->nE = nE;
    j->E.assign(energies, energies + nE);
    int rc = create(j);
    if (rc != QSIM_OK) {
        destroy(j);
        return rc;
    }
    *job = j;
    return QSIM_OK;
}

int qsim_hip_born_iterate(qsim_born_job *j, int max_iter, double tol, int *iters,
                          double *residual, qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_born_iterate");
    QSIM_CHECK(j, "null job");
    QSIM_CHECK(max_iter >= 0, "max_iter must be >= 0");
    hipStream_t s = as_stream(stream);
    for (int e = 0; e < j->nE; ++e) j->iters[e] = 0;
    if (max_iter > 0) {
        if (j->fused) QSIM_TRY(fused_iterate(j, max_iter, tol, s));
        else QSIM_TRY(unfused_iterate(j, max_iter, tol, s));
    }
    QSIM_HIP(hipStreamSynchronize(s));
    for (int e = 0; e < j->nE; ++e) {
        if (iters) iters[e] = j->iters[e];
        if (residual) residual[e] = j->resid[e];
    }
    return QSIM_OK;
}

int qsim_hip_born_get_dpsi(qsim_born_job *j, int e, qsim_c128 *out, qsim_stream_t stream)
{
    QSIM_CHECK(j && out, "null argument");
    QSIM_CHECK(e >= 0 && e < j->nE, "energy index out of range");
    hipStream_t s = as_stream(stream);
    c128 *o = reinterpret_cast<c128 *>(out);
    if (j->prec == 64) {
        QSIM_HIP(hipMemcpyAsync(o, j->host[e].dpsi, j->n * sizeof(c128), hipMemcpyDeviceToDevice, s));
    } else {
        const c64 *src = (const c64 *)j->host[e].dpsi;
        QSIM_TRY(each_b(j->n, s, [=] __device__(size_t i) { o[i] = cvt<double>(src[i]); }));
    }
    return QSIM_OK;
}

int qsim_hip_born_get_veff(qsim_born_job *j, int e, qsim_c128 *out, double *eps,
                           qsim_stream_t stream)
{
    QSIM_CHECK(j && out, "null argument");
    QSIM_CHECK(e >= 0 && e < j->nE, "energy index out of range");
    const BornDev J = j->host[e];
    const BornGeom geo = j->geo;
    const c128 *V = j->V;
    const int pre = j->pre;
    c128 *o = reinterpret_cast<c128 *>(out);
    QSIM_TRY(each_b(j->n, as_stream(stream), [=] __device__(size_t idx) {
        size_t i, jj, k;
        if (geo.dim == 3) {
            k = idx % geo.n2;
            jj = (idx / geo.n2) % geo.n1;
            i = idx / (geo.n1 * geo.n2);
        } else {
            k = 0;
            jj = idx % geo.n1;
            i = idx / geo.n1;
        }
        o[idx] = pre == PRE_NAIVE ? V[idx] : v_eff(V[idx], J, geo.dim, i, jj, k);
    }));
    if (eps) *eps = J.eps;
    return QSIM_OK;
}

int qsim_hip_born_xsection(qsim_born_job *j, int e, double cosx, double cosy, double cosz,
                           double *dsigma, qsim_c128 *amp, qsim_stream_t stream)
{
    QSIM_CHECK(j, "null job");
    QSIM_CHECK(e >= 0 && e < j->nE, "energy index out of range");
    const qsim_born_desc &d = j->d;
    const BornGeom geo = j->geo;
    const double En = j->E[e];
    const double k0 = std::sqrt(2 * En * d.mass) / d.hbar;
    double dx[3] = {1, 1, 1};
    for (int a = 0; a < d.dim; ++a) dx[a] = (d.hi[a] - d.lo[a]) / d.n[a];
    // r = sum (psi0 + dpsi) Re V e^{-i k.x}: the Xsection3D / 2D host-loop
    // sums (src/DeviceTemplate.h:1646-1665, src/ScatteringSolverImpl.cpp:
    // 186-191) with psi0 taken from the separable tables
    const BornDev J = j->host[e];
    const int prec = j->prec;
    const c128 *V = j->V;
    hipStream_t s = as_stream(stream);
    const double x0 = d.lo[0], y0 = d.lo[1], z0 = d.lo[2];
    const double ddx = dx[0], ddy = dx[1], ddz = dx[2];
    auto tot_at = [=] __device__(size_t idx, size_t i, size_t jj, size_t k) {
        const c128 dp = prec == 64 ? ((const c128 *)J.dpsi)[idx]
                                   : cvt<double>(((const c64 *)J.dpsi)[idx]);
        return psi0_at(J, geo.dim, i, jj, k) + dp;
    };
    double rs[2];
    if (d.dim == 3) {
        const double nm = 1 / std::sqrt(cosx * cosx + cosy * cosy + cosz * cosz);
        const double kx = k0 * cosx * nm, ky = k0 * cosy * nm, kz = k0 * cosz * nm;
        const size_t ny = geo.n1, nz = geo.n2;
        QSIM_TRY(red::reduce_sum<2>(j->n, s, rs, [=] __device__(size_t idx, double *v) {
            const size_t iz = idx % nz, q = idx / nz, iy = q % ny, ix = q / ny;
            const double ph = (x0 + (double)ix * ddx) * kx + (y0 + (double)iy * ddy) * ky +
                              (z0 + (double)iz * ddz) * kz;
            double sn, cs;
            sincos(-ph, &sn, &cs);
            const c128 t = (tot_at(idx, ix, iy, iz) * V[idx].x) * mk(cs, sn);
            v[0] = t.x;
            v[1] = t.y;
        }));
    } else {
        const double nm = 1 / std::sqrt(cosx * cosx + cosy * cosy);
        const double kx = k0 * cosx * nm, ky = k0 * cosy * nm;
        const size_t ny = geo.n1;
        QSIM_TRY(red::reduce_sum<2>(j->n, s, rs, [=] __device__(size_t idx, double *v) {
            const size_t ix = idx / ny, iy = idx % ny;
            const double x = x0 + ddx * (double)ix, y = y0 + ddy * (double)iy;
            const double ph = kx * x + ky * y;
            double sn, cs;
            sincos(ph, &sn, &cs);
            const c128 t = (tot_at(idx, ix, iy, 0) * V[idx].x) * mk(cs, -sn);
            v[0] = t.x;
            v[1] = t.y;
        }));
    }
    const qsim_c128 r{rs[0], rs[1]};
    const cd rr(r.re, r.im);
    double ds;
    if (d.dim == 3) {
        // src/ScatteringSolverImpl.cpp:355-357
        ds = std::norm(rr * (2 * d.mass) / (d.hbar * d.hbar) * dx[0] * dx[1] * dx[2] / (4 * M_PI));
    } else {
        // src/ScatteringSolverImpl.cpp:193
        ds = std::norm(rr * (2 * d.mass) / (d.hbar * d.hbar) * dx[0] * dx[1] * (1. / 4)) *
             (2 / (M_PI * k0));
    }
    if (dsigma) *dsigma = ds;
    if (amp) *amp = r;
    return QSIM_OK;
}

int qsim_hip_born_destroy(qsim_born_job *j)
{
    destroy(j);
    return QSIM_OK;
}

}  // extern "C"
