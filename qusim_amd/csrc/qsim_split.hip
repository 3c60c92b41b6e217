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
#include <algorithm>
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

// ---------------------------------------------- x axis split across rows
// With nx = R * M (R = 2 or 4) the x transform is taken in two stages (the
// four-step factorisation, x = M i1 + i2, kx = k1 + R k2):
//   A : y[k1][i2] = w_nx^(i2 k1) * sum_i1 x[M i1 + i2] w_R^(i1 k1)
//   B : X[k1 + R k2] = sum_i2 y[k1][i2] w_M^(i2 k2)
// and the inverse as B' then A'.  The row sweep does A' and A on the R rows
// {i2 + M k} of one group, which sit in R adjacent lanes (lane = R t + k), so
// the R-point transforms are lane exchanges (DPP quad permutations, no LDS);
// the column sweep then runs B, the kinetic factor and B' on R * ny lines of
// length M instead of ny lines of length nx.  Between sweeps, row k1 M + i2
// of the working array holds y[k1][i2]; at call boundaries psi is in natural
// order.
// DPP quad permutations: lane ^ 1 and lane ^ 2.
constexpr int kDppXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}

// One radix-2 butterfly across lanes lane and lane ^ m: the lane with the
// bit clear takes mine + partner, the other partner - mine (one rounding each).
template <int CTRL, typename T>
__device__ __forceinline__ cx<T> xbfly(cx<T> a, bool upper)
{
    const cx<T> b = mk(dpp_mov<CTRL>(a.x), dpp_mov<CTRL>(a.y));
    const T sg = upper ? T(-1) : T(1);  // sg * a is exact: one rounding per sum
    return mk(b.x + sg * a.x, b.y + sg * a.y);
}

// R-point DFT across the R lanes of a group, element by element.
// FWD: input in bit-reversed lane order, output natural (lane k holds X[k]).
// !FWD (inverse, +i): input natural, output bit-reversed.
template <int R, bool FWD, typename T, int E>
__device__ __forceinline__ void xrow_dft(cx<T> (&v)[E], int k)
{
    if constexpr (R == 2) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = xbfly<kDppXor1>(v[e], k & 1);
    } else {
        static_assert(R == 4, "row groups of 2 or 4");
        const bool b0 = k & 1, b1 = (k >> 1) & 1;
        if constexpr (FWD) {
            // over bit 0, then w4^(bit 0) on lanes with bit 1 set, over bit 1
#pragma unroll
            for (int e = 0; e < E; ++e) {
                cx<T> a = xbfly<kDppXor1>(v[e], b0);
                if (b1 && b0) a = mk(a.y, -a.x);  // * (-i)
                v[e] = xbfly<kDppXor2>(a, b1);
            }
        } else {
            // over bit 1, then w4^-(bit 1) on lanes with bit 0 set, over bit 0
#pragma unroll
            for (int e = 0; e < E; ++e) {
                cx<T> a = xbfly<kDppXor2>(v[e], b1);
                if (b0 && b1) a = mk(-a.y, a.x);  // * (+i)
                v[e] = xbfly<kDppXor1>(a, b0);
            }
        }
    }
}

// Row sweep of the x-split schedule: one group of R rows {i2 + M k} per
// workgroup (R * P::TPL threads, lane R t + k on row k of the group).
// EVP: load the MID sweep's phase factors before the inverse transform.
template <typename T, typename P, int R, int LSTR, int MODE, bool NT, bool EVP>
__device__ __forceinline__ void row_group_x(const SplitDev &J, size_t i2, size_t nx, int iev,
                                            const c128 *__restrict__ tw,
                                            const c128 *__restrict__ twx, cx<T> *lds)
{
    constexpr int N = P::TPL * P::E;
    const int tid = threadIdx.x;
    const int k = tid % R, t = tid / R;
    const size_t M = nx / R;
    const int kr = R == 4 ? ((k & 1) << 1) | (k >> 1) : k;  // bit-reversed k
    // natural-order row of lane k (holds x[M kr + i2]) and working-array row
    // (holds y[k][i2])
    const size_t xrow = M * kr + i2, yrow = M * k + i2;
    const c128 *__restrict__ ev = J.ev[iev];
    cx<T> *work = reinterpret_cast<cx<T> *>(J.work);
    c128 *psi = J.psi;
    const c128 wk = ld_g(twx + i2 * k);  // w_nx^(i2 k)
    constexpr int RA = P::RA_, RZ = P::RLAST;
    cx<T> v[P::E];
    if constexpr (MODE == SW_FIRST) {
        const size_t base = xrow * N;
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r) {
                const int pos = P::template in_pos<RA>(t, u, r);
                v[u * RA + r] = apply(conv<T>(ld_g(psi + base + pos)), ld_g(ev + base + pos));
            }
    } else {
        const size_t ybase = yrow * N, xbase = xrow * N;
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                v[u * RZ + r] = ld_line<NT>(work + ybase + P::template in_pos<RZ>(t, u, r));
        constexpr bool PREFETCH = EVP;
        c128 evp[PREFETCH ? P::E : 1];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r)
                    evp[u * RA + r] = ld_g(ev + xbase + P::template in_pos<RA>(t, u, r));
        }
        // A': w_nx^-(i2 k1), then the inverse R-point transform across rows
#pragma unroll
        for (int e = 0; e < P::E; ++e) v[e] = apply(v[e], mk(wk.x, -wk.y));
        xrow_dft<R, false>(v, k);
        P::inverse(v, t, lds + k * LSTR, tw);
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r) {
                const int pos = P::template in_pos<RA>(t, u, r);
                if constexpr (PREFETCH) v[u * RA + r] = apply(v[u * RA + r], evp[u * RA + r]);
                else v[u * RA + r] = apply(v[u * RA + r], ld_g(ev + xbase + pos));
            }
    }
    if constexpr (MODE == SW_LAST) {
        const size_t base = xrow * N;
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r)
                store_nt(&psi[base + P::template in_pos<RA>(t, u, r)], conv<double>(v[u * RA + r]));
    } else {
        P::forward(v, t, lds + k * LSTR, tw);
        // A: the forward R-point transform across rows, then w_nx^(i2 k1)
        xrow_dft<R, true>(v, k);
#pragma unroll
        for (int e = 0; e < P::E; ++e) v[e] = apply(v[e], wk);
        const size_t base = yrow * N;
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                store_nt(&work[base + P::template in_pos<RZ>(t, u, r)], v[u * RZ + r]);
    }
}

// The R-point transforms of xrow_dft with the R rows of a group in one
// thread (v[k] is what lane k holds there): the same butterflies, partner
// first, so the results are those of the lane exchanges bit for bit.
template <typename T>
__device__ __forceinline__ cx<T> lbfly(cx<T> partner, cx<T> mine, bool upper)
{
    const T sg = upper ? T(-1) : T(1);
    return mk(partner.x + sg * mine.x, partner.y + sg * mine.y);
}

template <int R, bool FWD, typename T, int E>
__device__ __forceinline__ void xrow_dft_local(cx<T> (&v)[R][E])
{
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if constexpr (R == 2) {
            const cx<T> a0 = v[0][e], a1 = v[1][e];
            v[0][e] = lbfly(a1, a0, false);
            v[1][e] = lbfly(a0, a1, true);
        } else {
            static_assert(R == 4, "row groups of 2 or 4");
            constexpr int M1 = FWD ? 1 : 2, M2 = FWD ? 2 : 1;
            cx<T> a[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) a[k] = lbfly(v[k ^ M1][e], v[k][e], (k & M1) != 0);
            a[3] = FWD ? mk(a[3].y, -a[3].x) : mk(-a[3].y, a[3].x);
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k][e] = lbfly(a[k ^ M2], a[k], (k & M2) != 0);
        }
    }
}

// Row sweep of the x-split schedule with the R rows of a group in each
// thread (P::TPL threads, thread t holds the line elements of t in every
// row): the cross-row transforms need no lane exchange, and the R line
// transforms run one after the other through one LDS line.  Rows of 4096
// points take this layout: one lane per row needs R * TPL threads with a
// half or a quarter of the register budget each, and those groups spilled
// to scratch.  Results equal those of row_group_x.
template <typename T, typename P, int R, int MODE, bool NT, bool EVP>
__device__ __forceinline__ void row_group_xl(const SplitDev &J, size_t i2, size_t nx, int iev,
                                             const c128 *__restrict__ tw,
                                             const c128 *__restrict__ twx, cx<T> *lds)
{
    constexpr int N = P::TPL * P::E;
    const int t = threadIdx.x;
    const size_t M = nx / R;
    const c128 *__restrict__ ev = J.ev[iev];
    cx<T> *work = reinterpret_cast<cx<T> *>(J.work);
    c128 *psi = J.psi;
    constexpr int RA = P::RA_, RZ = P::RLAST;
    // v[k]: the row of lane k in row_group_x (natural row M bitrev(k) + i2,
    // working-array row M k + i2)
    cx<T> v[R][P::E];
    c128 wk[R];
#pragma unroll
    for (int k = 0; k < R; ++k) wk[k] = ld_g(twx + i2 * k);
    auto xrow = [&](int k) {
        const int kr = R == 4 ? ((k & 1) << 1) | (k >> 1) : k;
        return M * kr + i2;
    };
    if constexpr (MODE == SW_FIRST) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const size_t base = xrow(k) * N;
#pragma unroll
            for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r) {
                    const int pos = P::template in_pos<RA>(t, u, r);
                    v[k][u * RA + r] =
                        apply(conv<T>(ld_g(psi + base + pos)), ld_g(ev + base + pos));
                }
        }
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const size_t ybase = (M * k + i2) * N;
#pragma unroll
            for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    v[k][u * RZ + r] =
                        ld_line<NT>(work + ybase + P::template in_pos<RZ>(t, u, r));
        }
        c128 evp[EVP ? R : 1][EVP ? P::E : 1];
        if constexpr (EVP) {
#pragma unroll
            for (int k = 0; k < R; ++k)
#pragma unroll
                for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                    for (int r = 0; r < RA; ++r)
                        evp[k][u * RA + r] =
                            ld_g(ev + xrow(k) * N + P::template in_pos<RA>(t, u, r));
        }
        // A': w_nx^-(i2 k1), then the inverse R-point transform across rows
#pragma unroll
        for (int k = 0; k < R; ++k)
#pragma unroll
            for (int e = 0; e < P::E; ++e) v[k][e] = apply(v[k][e], mk(wk[k].x, -wk[k].y));
        xrow_dft_local<R, false>(v);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            P::inverse(v[k], t, lds, tw);
            const size_t xbase = xrow(k) * N;
#pragma unroll
            for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r) {
                    const int pos = P::template in_pos<RA>(t, u, r);
                    if constexpr (EVP) v[k][u * RA + r] = apply(v[k][u * RA + r], evp[k][u * RA + r]);
                    else v[k][u * RA + r] = apply(v[k][u * RA + r], ld_g(ev + xbase + pos));
                }
        }
    }
    if constexpr (MODE == SW_LAST) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const size_t base = xrow(k) * N;
#pragma unroll
            for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r)
                    store_nt(&psi[base + P::template in_pos<RA>(t, u, r)],
                             conv<double>(v[k][u * RA + r]));
        }
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) P::forward(v[k], t, lds, tw);
        // A: the forward R-point transform across rows, then w_nx^(i2 k1)
        xrow_dft_local<R, true>(v);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const size_t base = (M * k + i2) * N;
#pragma unroll
            for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    store_nt(&work[base + P::template in_pos<RZ>(t, u, r)],
                             apply(v[k][u * RZ + r], wk[k]));
        }
    }
}

// Launch shape of the x-split row sweep: R lines of Line P per workgroup,
// lanes line-fast within each quad.  The LDS line stride (mod 16) is the
// measured best of the exchanges, as for the line-fast column tiles: 16/R
// for 8-byte elements, 8/R for 16-byte ones (FP64 1024-point rows, R = 2:
// 2.17 -> 1.50 times the conflict-free LDS cycles).
template <typename T, int N, int V, int R>
struct XRowLaunch {
    // rows of 4096 points: the R rows in each thread (row_group_xl; FP32
    // 8192-point rows spilled more in that layout than one lane per row)
    static constexpr bool LOCAL = N == 4096;
    // FP64 groups of rows, lane-interleaved, take the m ^ ((m >> 2) & 7)
    // swizzle (1.50 -> 1.11 times the conflict-free LDS cycles at 1024
    // points, R = 2); FP32 groups and the in-thread layout keep the padding
    using Line = typename P2Cfg<T, N, V>::Line::template WithSwz<
        (!LOCAL && sizeof(cx<T>) == 16) ? 2 : 0>;
    static constexpr int THREADS = LOCAL ? Line::TPL : R * Line::TPL;
    static constexpr int TARGET = (sizeof(cx<T>) == 8 ? 16 : 8) / R;
    static constexpr int LSTR =
        LOCAL ? Line::LS : Line::LS_RAW + (((TARGET - Line::LS_RAW % 16) % 16) + 16) % 16;
    static constexpr size_t LDS_BYTES = (size_t)(LOCAL ? 1 : R) * LSTR * sizeof(cx<T>);
    static constexpr bool OK = THREADS <= 1024 && LDS_BYTES <= 160 * 1024;
    // The MID sweep prefetches its phase factors (4 registers each) next to
    // the line elements (4 or 2 registers each) only where those take at
    // most a quarter of the per-lane register budget of THREADS-thread
    // workgroups; beyond that the 4096-point groups spilled to scratch.
    static constexpr int BUDGET = 512 / cmax(1, (THREADS / 64 + 3) / 4);
    static constexpr int NEED = (LOCAL ? R : 1) * Line::E * (4 + 2 * (int)(sizeof(T) / 4));
    static constexpr bool EVP = Line::E <= 16 && 4 * NEED <= BUDGET;
};

template <typename T, int N, int MODE, int V, bool NT, int R>
__global__ void __launch_bounds__((XRowLaunch<T, N, V, R>::THREADS))
    k_row_sweep_x(const SplitDev *__restrict__ jobs, size_t nx, int iev,
                  const c128 *__restrict__ tw, const c128 *__restrict__ twx)
{
    using XL = XRowLaunch<T, N, V, R>;
    extern __shared__ unsigned char smem_raw[];
    if constexpr (XL::LOCAL)
        row_group_xl<T, typename XL::Line, R, MODE, NT, XL::EVP>(
            jobs[blockIdx.y], blockIdx.x, nx, iev, tw, twx, reinterpret_cast<cx<T> *>(smem_raw));
    else
        row_group_x<T, typename XL::Line, R, XL::LSTR, MODE, NT, XL::EVP>(
            jobs[blockIdx.y], blockIdx.x, nx, iev, tw, twx, reinterpret_cast<cx<T> *>(smem_raw));
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

// Column sweep over the L lines g0 .. g0 + L - 1 of one job, lanes
// line-fast (lane i on line i % L).  R = 1: line g is column g of the
// nx(=M) x ncols array.  R > 1 (x split, nx = R M): line g is (k1, c) =
// (g / ncols, g % ncols), the M elements at rows k1 M + i2 of column c, and
// its output position k2 is kx = k1 + R k2 of the kinetic factor.
template <typename T, typename P, int L, int LSTR, bool NT, int R = 1>
__device__ __forceinline__ void col_lines(const SplitDev &J, size_t g0, size_t ncols, int ik,
                                          const c128 *__restrict__ tw, cx<T> *lds)
{
    constexpr int M = P::TPL * P::E;
    const int tid = threadIdx.x;
    const int l = tid % L, t = tid / L;
    const size_t g = g0 + l;
    const bool active = g < (size_t)R * ncols;
    const size_t ga = active ? g : 0;
    const size_t k1 = R == 1 ? 0 : ga / ncols, c = R == 1 ? ga : ga % ncols;
    const size_t base = k1 * M * ncols + c;
    cx<T> *work = reinterpret_cast<cx<T> *>(J.work);
    const c128 *__restrict__ kx = J.kx[ik] + k1;
    const c128 kyj = ld_g(J.ky[ik] + c);
    constexpr int RA = P::RA_, RZ = P::RLAST;
    cx<T> v[P::E];
#pragma unroll
    for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
        for (int r = 0; r < RA; ++r)
            v[u * RA + r] =
                active ? ld_line<NT>(work + base + (size_t)P::template in_pos<RA>(t, u, r) * ncols)
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
                kxp[u * RZ + r] = ld_g(kx + (size_t)R * P::template in_pos<RZ>(t, u, r));
    }
    P::forward(v, t, lds + l * LSTR, tw, l);
#pragma unroll
    for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const c128 kk = PREFETCH ? kxp[u * RZ + r]
                                     : ld_g(kx + (size_t)R * P::template in_pos<RZ>(t, u, r));
            v[u * RZ + r] = apply(v[u * RZ + r], kk * kyj);
        }
    P::inverse(v, t, lds + l * LSTR, tw, l);
    if (active) {
#pragma unroll
        for (int u = 0; u < P::E / RA; ++u)
#pragma unroll
            for (int r = 0; r < RA; ++r)
                store_nt(&work[base + (size_t)P::template in_pos<RA>(t, u, r) * ncols],
                         v[u * RA + r]);
    }
}

template <typename T, int N, int V, bool NT, int R = 1>
__global__ void __launch_bounds__((P2Launch<T, N, true, V>::THREADS))
    k_col_sweep(const SplitDev *__restrict__ jobs, size_t ncols, int ik,
                const c128 *__restrict__ tw)
{
    using LP = P2Launch<T, N, true, V>;
    extern __shared__ unsigned char smem_raw[];
    col_lines<T, typename LP::Line, LP::L, LP::LSTR, NT, R>(
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
    // x-split factor R (1, 2, 4): the column sweep runs R * ny lines of
    // nx / R points, the row sweep the cross-row stages (tw_xm: twiddles of
    // nx / R)
    int xsplit = 1;
    const void *tw_xm = nullptr;
    // replay captured sweep sequences (keyed by step count, oldest first in
    // graph_order)
    bool use_graph = false;
    std::map<size_t, hipGraphExec_t> graphs;
    hipEvent_t graph_done = nullptr;  // follows the job's last replay
    bool replayed = false;
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

// x-split row sweep (R rows per workgroup); `dry` only checks that the row
// group fits one workgroup (threads, LDS)
template <typename T, int N, int MODE, int V, int R>
int launch_row_x(const SplitDev *jobs, int njobs, size_t nx, int iev, const void *tw,
                 const void *twx, hipStream_t s, bool dry)
{
    using XL = XRowLaunch<T, N, V, R>;
    if constexpr (!XL::OK) {
        return set_error(QSIM_ERR_UNSUPPORTED, "x split: the row group exceeds one workgroup");
    } else {
        if (dry) return QSIM_OK;
        auto fn = k_row_sweep_x<T, N, MODE, V, false, R>;
        if constexpr (N <= 1024)
            if (line_nt(N, (size_t)njobs * nx * N * sizeof(cx<T>)))
                fn = k_row_sweep_x<T, N, MODE, V, true, R>;
        QSIM_TRY(lds_limit((const void *)fn, XL::LDS_BYTES));
        dim3 grid((unsigned)(nx / R), (unsigned)njobs);
        hipLaunchKernelGGL(fn, grid, dim3(XL::THREADS), XL::LDS_BYTES, s, jobs, nx, iev,
                           (const c128 *)tw, (const c128 *)twx);
        QSIM_HIP(hipGetLastError());
        return QSIM_OK;
    }
}

template <typename T, int N, int V, int R>
int launch_col(const SplitDev *jobs, int njobs, size_t ncols, int ik, const void *tw,
               hipStream_t s)
{
    using LP = P2Launch<T, N, true, V>;
    auto fn = k_col_sweep<T, N, V, false, R>;
    if constexpr (N <= 1024)
        if (line_nt(N, (size_t)njobs * R * ncols * N * sizeof(cx<T>)))
            fn = k_col_sweep<T, N, V, true, R>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    dim3 grid((unsigned)((R * ncols + LP::L - 1) / LP::L), (unsigned)njobs);
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

// Launch variants: lengths 1024 and 2048 have the candidates 0..7, lengths
// 128, 256, 512 and 4096 the two radix plans 0 (16 or 32 elements per
// thread) and 1 (8 per thread), every other length only variant 0 (P2Cfg).
constexpr int kVariants = 8;

int num_variants(size_t n)
{
    return (n == 1024 || n == 2048) ? kVariants
           : (n == 128 || n == 256 || n == 512 || n == 4096) ? 2
                                                             : 1;
}
bool has_variants(size_t n) { return num_variants(n) > 1; }

// Variants a new job starts with: the pairs the autotuner picked most often
// on MI355X for single and batched jobs of these lengths (row variant by the
// row length ny, column variant by the column length nx).  Fixed, so jobs that
// are never tuned (request loops, batches) still take a fast pair and two
// runs of one job always choose the same radix plans.
// 128- to 512-point lines take the 8-per-thread plan (variant 1): measured
// on MI355X per step, single job / 16 batched: 128^2 FP32 11.7 -> 9.6 /
// 13.4 -> 11.2 us, 256^2 FP64 13.7 -> 11.0 / 23.9 -> 20.9 us, 512^2 FP32
// 25.9 -> 16.1 / 49.0 -> 51.4 us.  (512 FP32 columns in plan 0 batch 13%
// faster, but their 32 elements per thread would set the register budget of
// the multi-size group kernels, which mirror these defaults.)
// 4096-point rows take plan 1 as well (4096^2 x-split rows FP64 / FP32:
// 187-191 / 128-135 us); 4096-point columns (no x-split) plan 1 in FP64
// (345 against 516 us) and plan 0 in FP32 (180 against 543 us).
__host__ __device__ constexpr int default_variant(int prec, size_t n, bool row)
{
    return n == 1024   ? (prec == 64 ? 4 : 7)
           : n == 2048 ? (prec == 64 ? (row ? 7 : 5) : (row ? 7 : 3))
           : n == 4096 ? (row || prec == 64 ? 1 : 0)
           : (n == 128 || n == 256 || n == 512) ? 1
                                                : 0;
}

// The x-split row sweep depends on the radix plan only (its lines per
// workgroup are the R rows of a group): variants with the same plan share one
// kernel.  1024: 0/1/5 (16-16-4), 2/3 (32-32), 4 (8-8-16), 6/7 (8-8-4-4);
// 2048: 0/1/4/5 (16-16-8), 2 (32-32-2), 3 (8-16-16), 6/7 (8-8-8-4).
constexpr int xrow_plan(int n, int var)
{
    if (n == 1024) return (var == 1 || var == 5) ? 0 : (var == 3 ? 2 : (var == 7 ? 6 : var));
    if (n == 2048) return (var == 1 || var == 4 || var == 5) ? 0 : (var == 7 ? 6 : var);
    return 0;
}

template <typename T, int NN, int MODE, int R>
int row_vx(int var, const SplitDev *jobs, int njobs, size_t nx, int iev, const void *tw,
           const void *twx, hipStream_t s, bool dry)
{
    if constexpr (NN == 128 || NN == 256 || NN == 512 || NN == 4096)
        if (var == 1)
            return launch_row_x<T, NN, MODE, 1, R>(jobs, njobs, nx, iev, tw, twx, s, dry);
    if constexpr (NN == 1024 || NN == 2048) {
        switch (xrow_plan(NN, var)) {
            case 2: return launch_row_x<T, NN, MODE, 2, R>(jobs, njobs, nx, iev, tw, twx, s, dry);
            case 3: return launch_row_x<T, NN, MODE, 3, R>(jobs, njobs, nx, iev, tw, twx, s, dry);
            case 4: return launch_row_x<T, NN, MODE, 4, R>(jobs, njobs, nx, iev, tw, twx, s, dry);
            case 6: return launch_row_x<T, NN, MODE, 6, R>(jobs, njobs, nx, iev, tw, twx, s, dry);
            default: break;
        }
    }
    return launch_row_x<T, NN, MODE, 0, R>(jobs, njobs, nx, iev, tw, twx, s, dry);
}

template <typename T, int NN, int MODE>
int row_v(int var, int R, const SplitDev *jobs, int njobs, size_t nrows, int iev, const void *tw,
          const void *twx, hipStream_t s, bool dry)
{
    if (R == 2) return row_vx<T, NN, MODE, 2>(var, jobs, njobs, nrows, iev, tw, twx, s, dry);
    if (R == 4) return row_vx<T, NN, MODE, 4>(var, jobs, njobs, nrows, iev, tw, twx, s, dry);
    if (dry) return QSIM_OK;
    if constexpr (NN == 128 || NN == 256 || NN == 512 || NN == 4096)
        if (var == 1) return launch_row<T, NN, MODE, 1>(jobs, njobs, nrows, iev, tw, s);
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

template <typename T, int NN, int R>
int col_vr(int var, const SplitDev *jobs, int njobs, size_t ncols, int ik, const void *tw,
           hipStream_t s)
{
    if constexpr (NN == 128 || NN == 256 || NN == 512 || NN == 4096)
        if (var == 1) return launch_col<T, NN, 1, R>(jobs, njobs, ncols, ik, tw, s);
    if constexpr (NN == 1024 || NN == 2048) {
        switch (var) {
            case 1: return launch_col<T, NN, 1, R>(jobs, njobs, ncols, ik, tw, s);
            case 2: return launch_col<T, NN, 2, R>(jobs, njobs, ncols, ik, tw, s);
            case 3: return launch_col<T, NN, 3, R>(jobs, njobs, ncols, ik, tw, s);
            case 4: return launch_col<T, NN, 4, R>(jobs, njobs, ncols, ik, tw, s);
            case 5: return launch_col<T, NN, 5, R>(jobs, njobs, ncols, ik, tw, s);
            case 6: return launch_col<T, NN, 6, R>(jobs, njobs, ncols, ik, tw, s);
            case 7: return launch_col<T, NN, 7, R>(jobs, njobs, ncols, ik, tw, s);
            default: break;
        }
    }
    return launch_col<T, NN, 0, R>(jobs, njobs, ncols, ik, tw, s);
}

template <typename T, int NN>
int col_v(int var, int R, const SplitDev *jobs, int njobs, size_t ncols, int ik, const void *tw,
          hipStream_t s)
{
    if (R == 2) return col_vr<T, NN, 2>(var, jobs, njobs, ncols, ik, tw, s);
    if (R == 4) return col_vr<T, NN, 4>(var, jobs, njobs, ncols, ik, tw, s);
    return col_vr<T, NN, 1>(var, jobs, njobs, ncols, ik, tw, s);
}

// Row sweep of rows of length ny (nrows = nx of them); R = x-split factor
// (twx: the length-nx twiddles of the cross-row stage).
int row_sweep(int prec, size_t ny, int mode, int var, int R, const SplitDev *jobs, int njobs,
              size_t nrows, int iev, const void *tw, const void *twx, hipStream_t s,
              bool dry = false)
{
#define ROWCALL(T, NN)                                                                        \
    (mode == SW_FIRST                                                                         \
         ? row_v<T, NN, SW_FIRST>(var, R, jobs, njobs, nrows, iev, tw, twx, s, dry)           \
         : mode == SW_MID ? row_v<T, NN, SW_MID>(var, R, jobs, njobs, nrows, iev, tw, twx, s, \
                                                 dry)                                         \
                          : row_v<T, NN, SW_LAST>(var, R, jobs, njobs, nrows, iev, tw, twx,   \
                                                  s, dry))
    QSIM_SWITCH_N(prec, ny, ROWCALL)
#undef ROWCALL
}

// Column sweep: R * ncols lines of length m = nx / R (tw: twiddles of m).
int col_sweep(int prec, size_t m, int var, int R, const SplitDev *jobs, int njobs, size_t ncols,
              int ik, const void *tw, hipStream_t s)
{
#define COLCALL(T, NN) col_v<T, NN>(var, R, jobs, njobs, ncols, ik, tw, s)
    QSIM_SWITCH_N(prec, m, COLCALL)
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

// x-split factor a new job starts with.
int default_xsplit(int prec, size_t nx)
{
    return nx == 4096 || (prec == 32 && nx == 8192) ? 2 : 1;
}

// x-split factors a fused 2D job can take: 1, or 2 / 4 when the column
// lines nx / R have a fused kernel and a row group of R rows fits one
// workgroup with row variant `rv` (threads, LDS).
int xsplit_check(const qsim_split_job *j, int R, int rv)
{
    if (R == 1) return QSIM_OK;
    if (R != 2 && R != 4) return set_error(QSIM_ERR_INVALID, "x split factor must be 1, 2 or 4");
    if (!j->fused || j->d.dim != 2)
        return set_error(QSIM_ERR_UNSUPPORTED, "x split needs a fused 2D job");
    const int prec = j->d.precision;
    if (j->d.nx % R || !fused_len_ok(j->d.nx / R, prec, 2))
        return set_error(QSIM_ERR_UNSUPPORTED, "x split: no fused column kernel for nx / R");
    return row_sweep(prec, j->d.ny, SW_MID, rv, R, nullptr, 0, j->d.nx, 0, nullptr, nullptr,
                     nullptr, true);
}

// Sets the x-split factor (twiddles of nx / R for the column lines); the
// column variant restarts from the default of the new line length.
int set_xsplit(qsim_split_job *j, int R, int rv)
{
    QSIM_TRY(xsplit_check(j, R, rv));
    if (R != 1) QSIM_TRY(twiddle_table(j->d.nx / R, j->d.precision, &j->tw_xm));
    if (R != j->xsplit) j->col_var = default_variant(j->d.precision, j->d.nx / R, false);
    j->xsplit = R;
    return QSIM_OK;
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
    const int rv = j->row_var, cv = j->col_var, R = j->xsplit;
    const void *twc = R == 1 ? j->tw_x : j->tw_xm;
    TIMED(tm, TK_ROW_FIRST,
          row_sweep(prec, ny, SW_FIRST, rv, R, jobs, njobs, nx, j->ev_start, j->tw_y, j->tw_x, s));
    for (size_t st = 0; st < nsteps; ++st) {
        for (int ph = 0; ph < j->nph; ++ph) {
            TIMED(tm, TK_COL,
                  col_sweep(prec, nx / R, cv, R, jobs, njobs, ny, j->ph_k[ph], twc, s));
            const bool last = (st + 1 == nsteps) && (ph + 1 == j->nph);
            if (last)
                TIMED(tm, TK_ROW_LAST, row_sweep(prec, ny, SW_LAST, rv, R, jobs, njobs, nx,
                                                 j->ev_end, j->tw_y, j->tw_x, s));
            else
                TIMED(tm, TK_ROW, row_sweep(prec, ny, SW_MID, rv, R, jobs, njobs, nx,
                                            j->ph_e[ph], j->tw_y, j->tw_x, s));
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

// Captured graphs may still be replaying: wait for the job's last replay
// (not for the whole device) before destroying them.
void wait_replays(qsim_split_job *j)
{
    if (j->replayed) (void)hipEventSynchronize(j->graph_done);
    j->replayed = false;
}

void drop_graphs(qsim_split_job *j)
{
    if (j->graphs.empty()) return;
    wait_replays(j);
    for (auto &kv : j->graphs) (void)hipGraphExecDestroy(kv.second);
    j->graphs.clear();
    j->graph_order.clear();
}

void destroy(qsim_split_job *j)
{
    if (!j) return;
    drop_graphs(j);
    if (j->graph_done) (void)hipEventDestroy(j->graph_done);
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
        if (d.dim == 2) {
            QSIM_TRY(twiddle_table(d.ny, d.precision, &j->tw_y));
            // columns too long for a tile of several lines in LDS (4096, FP32
            // 8192) start x-split in two: measured on MI355X at 4096^2, the
            // step went from 548 to 375 us (FP64: columns 345 -> 176 us) and
            // from 339 to 270 us (FP32-mixed)
            const int R = default_xsplit(d.precision, d.nx);
            if (R != 1 && xsplit_check(j, R, j->row_var) == QSIM_OK)
                QSIM_TRY(set_xsplit(j, R, j->row_var));
        }
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
    QSIM_DEVICE_WAIT_GUARD;
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
    std::unique_lock<std::recursive_mutex> lk(device_wait_lock(), std::defer_lock);
    if (it == j->graphs.end()) lk.lock();  // eviction and capture
    if (it == j->graphs.end() && j->graphs.size() >= kMaxGraphs) {
        // a replay of the graph being dropped may still be in flight
        wait_replays(j);
        auto old = j->graphs.find(j->graph_order.front());
        (void)hipGraphExecDestroy(old->second);
        j->graphs.erase(old);
        j->graph_order.erase(j->graph_order.begin());
    }
    if (it == j->graphs.end()) {
        if (!j->graph_done)
            QSIM_HIP(hipEventCreateWithFlags(&j->graph_done, hipEventDisableTiming));
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
    // replays on another stream than the last one follow it, so graph_done
    // always covers every replay in flight
    if (j->replayed) QSIM_HIP(hipStreamWaitEvent(s, j->graph_done, 0));
    QSIM_HIP(hipGraphLaunch(it->second, s));
    QSIM_HIP(hipEventRecord(j->graph_done, s));
    j->replayed = true;
    return QSIM_OK;
}

int qsim_hip_split_steps_timed(qsim_split_job *j, size_t nsteps, double *kernel_ms,
                               int *kernel_launches, qsim_stream_t stream)
{
    QSIM_DEVICE_WAIT_GUARD;
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
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(j, "null job");
    j->use_graph = enable != 0;
    if (!j->use_graph) drop_graphs(j);
    return QSIM_OK;
}

int qsim_hip_split_set_variants(qsim_split_job *j, int row_variant, int col_variant)
{
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(j, "null job");
    QSIM_CHECK(row_variant >= 0 && row_variant < kVariants && col_variant >= 0 &&
                   col_variant < kVariants,
               "launch variant out of range");
    const int rv = row_variant < num_variants(j->d.ny) ? row_variant : 0;
    if (j->fused && j->d.dim == 2) QSIM_TRY(xsplit_check(j, j->xsplit, rv));
    j->row_var = rv;
    j->col_var = col_variant < num_variants(j->d.nx / j->xsplit) ? col_variant : 0;
    drop_graphs(j);
    return QSIM_OK;
}

int qsim_hip_split_set_xsplit(qsim_split_job *j, int xsplit)
{
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(j, "null job");
    QSIM_TRY(set_xsplit(j, xsplit, j->row_var));
    drop_graphs(j);
    return QSIM_OK;
}

int qsim_hip_split_get_schedule(const qsim_split_job *j, int *row_variant, int *col_variant,
                                int *xsplit)
{
    QSIM_CHECK(j, "null job");
    if (row_variant) *row_variant = j->row_var;
    if (col_variant) *col_variant = j->col_var;
    if (xsplit) *xsplit = j->xsplit;
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
    // (row, column) pairs.  For each x-split factor R the job can take: the
    // row variant with the column variant fixed, then the column variant
    // with the chosen row variant; the fastest (R, row, column) wins.
    int bestR = j->xsplit, bestv[2] = {j->row_var, j->col_var};
    double bestpair = 1e30, bestus[2] = {0, 0};
    for (int R = 1; R <= 4 && rc == QSIM_OK; R *= 2) {
        const size_t m = j->d.nx / R;
        if (xsplit_check(j, R, 0) != QSIM_OK && xsplit_check(j, R, default_variant(prec, j->d.ny, true)) != QSIM_OK)
            continue;
        const void *twc = j->tw_x;
        if (R != 1 && twiddle_table(m, prec, &twc) != QSIM_OK) continue;
        int var[2] = {R == j->xsplit ? j->row_var : default_variant(prec, j->d.ny, true),
                      R == j->xsplit ? j->col_var : default_variant(prec, m, false)};
        if (xsplit_check(j, R, var[0]) != QSIM_OK) var[0] = 0;
        double best[2] = {1e30, 1e30};
        auto pair = [&](int vr_, int vc_) {
            int r2 = row_sweep(prec, j->d.ny, SW_MID, vr_, R, jt, njobs, j->d.nx, j->ph_e[0],
                               j->tw_y, j->tw_x, s);
            if (r2 == QSIM_OK)
                r2 = col_sweep(prec, m, vc_, R, jt, njobs, j->d.ny, j->ph_k[0], twc, s);
            return r2;
        };
        for (int which = 0; which < 2 && rc == QSIM_OK; ++which) {
            const size_t nlen = which == 0 ? j->d.ny : m;
            const int nv = num_variants(nlen);
            int chosen = var[which];
            for (int v = 0; v < nv && rc == QSIM_OK; ++v) {
                const int vr_ = which == 0 ? v : var[0], vc_ = which == 1 ? v : var[1];
                if (which == 0 && xsplit_check(j, R, vr_) != QSIM_OK) continue;
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
        if (rc == QSIM_OK && best[1] < bestpair) {
            bestpair = best[1];
            bestR = R;
            bestv[0] = var[0];
            bestv[1] = var[1];
            bestus[0] = best[0];
            bestus[1] = best[1];
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(s);
    release();
    if (rc != QSIM_OK) return rc;
    for (int i = 0; i < njobs; ++i) {
        QSIM_TRY(set_xsplit(jobs[i], bestR, bestv[0]));
        jobs[i]->row_var = bestv[0];
        jobs[i]->col_var = bestv[1];
        drop_graphs(jobs[i]);
    }
    if (row_variant) *row_variant = bestv[0];
    if (col_variant) *col_variant = bestv[1];
    if (row_us) *row_us = bestus[0];
    if (col_us) *col_us = bestus[1];
    return QSIM_OK;
}

}  // namespace

extern "C" {

int qsim_hip_split_autotune(qsim_split_job *j, int reps, int *row_variant, int *col_variant,
                            double *row_us, double *col_us, qsim_stream_t stream)
{
    QSIM_DEVICE_WAIT_GUARD;
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
    QSIM_DEVICE_WAIT_GUARD;
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
    QSIM_DEVICE_WAIT_GUARD;
    // every qsim_hip_split_steps call starts from the caller's psi
    QSIM_CHECK(j, "null job");
    return QSIM_OK;
}

int qsim_hip_split_norm2(qsim_split_job *j, double *out, qsim_stream_t s)
{
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(j && out, "null argument");
    return qsim_hip_norm2(reinterpret_cast<qsim_c128 *>(j->psi), j->n, out, s);
}

int qsim_hip_split_destroy(qsim_split_job *j)
{
    QSIM_DEVICE_WAIT_GUARD;
    destroy(j);
    return QSIM_OK;
}

int qsim_hip_split_steps_batched(qsim_split_job **jobs, int njobs, size_t nsteps,
                                 qsim_stream_t stream)
{
    QSIM_DEVICE_WAIT_GUARD;
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
               ji->d.precision == j0->d.precision && ji->row_var == j0->row_var &&
               ji->col_var == j0->col_var && ji->xsplit == j0->xsplit;
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
    QSIM_DEVICE_WAIT_GUARD;
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
                       ji->col_var == j0->col_var && ji->xsplit == j0->xsplit,
                   "batch: jobs differ in shape, method, precision or launch schedule");
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
    QSIM_DEVICE_WAIT_GUARD;
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
               ji->row_var == j0->row_var && ji->col_var == j0->col_var &&
               ji->xsplit == j0->xsplit;
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

// ------------------------------------------------------ multi-size job groups
// A group advances fused 2D split jobs of different grid sizes together.
// Jobs are dealt into length classes (the length of the lines a sweep
// transforms: ny for row sweeps, nx for column sweeps); each class has a
// kernel of its own (the register allocation of the per-length kernels) that
// loops, from a fixed grid, over the class's part of a device-side list of
// (job, tile) items.  Each job keeps its own step program: the row pass of
// global index P runs job j's FIRST sweep when P == first_pass(j), its LAST
// when P == first_pass(j) + nph * nsteps(j), a MID sweep in between and
// nothing otherwise; the column pass P runs while first_pass(j) <= P <
// first_pass(j) + nph * nsteps(j).  So every job sees exactly the sweep
// sequence of qsim_hip_split_steps(job, nsteps), in the same radix plans
// (its default launch variants), and its result is the same bit for bit,
// whatever else shares the launches.  An advance runs, for each shape class
// (nx, ny), its chain of passes; the chains are independent, so a captured
// advance forks them onto parallel branches and joins them before a final
// kernel adds the number of passes to the device-side pass index.  Work
// lists, counts and the pass index live in device memory: a captured advance
// stays valid while jobs of the same shapes join and leave.
namespace qsim {

constexpr int kGroupLens = 6;      // line lengths 64 .. 2048
constexpr int kGroupShapes = kGroupLens * kGroupLens;

struct GroupJob {
    SplitDev d;
    const c128 *twx, *twy;
    unsigned nx, ny;
    int nph, ev_start, ev_end;
    int ph_k[3], ph_e[3];
    long long first_pass, nsteps;
};

struct GroupItem {
    int job;        // -1: padding
    unsigned line0; // first row (row pass) or column (column pass) of the tile
};

// Per shape class s (nx, ny): its row items at items[off[4s] ..), count
// off[4s+1]; its column items at items[off[4s+2] ..), count off[4s+3] (a
// multiple of 8).
struct GroupCounts {
    int off[4 * kGroupShapes];
};

// One length's launch shape in the group kernels: that of the standalone
// sweep kernel in the job's default variant (radix plan, lines per
// workgroup, LDS strides).
template <typename T, int N, bool LF>
struct GroupLaunch : P2Launch<T, N, LF, default_variant(sizeof(T) == 8 ? 64 : 32, N, !LF)> {
};

__host__ __device__ constexpr bool group_len_ok(size_t n)
{
    return n >= 64 && n <= 2048 && (n & (n - 1)) == 0;
}
__host__ __device__ constexpr int group_len_index(size_t n)
{
    return n <= 64 ? 0 : 1 + group_len_index(n / 2);
}

template <typename T, int N, bool NT>
__global__ void __launch_bounds__((GroupLaunch<T, N, false>::THREADS))
    k_row_group(const GroupJob *__restrict__ jobs, const GroupItem *__restrict__ items,
                const GroupCounts *__restrict__ counts, const long long *__restrict__ basep,
                int r, int shape)
{
    using GL = GroupLaunch<T, N, false>;
    using P = typename GL::Line;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const long long Pidx = *basep + r;
    const GroupItem *list = items + counts->off[4 * shape];
    const int nitems = counts->off[4 * shape + 1];
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const GroupItem it = list[w];
        if (it.job < 0) continue;
        const GroupJob &G = jobs[it.job];
        const long long q = Pidx - G.first_pass, qend = (long long)G.nph * G.nsteps;
        if (q < 0 || q > qend) continue;  // the job is not running at this pass
        if (q == 0) {
            row_lines<T, P, GL::L, GL::LSTR, SW_FIRST, NT>(G.d, it.line0, G.nx, G.ev_start, G.twy,
                                                           lds);
        } else if (q == qend) {
            row_lines<T, P, GL::L, GL::LSTR, SW_LAST, NT>(G.d, it.line0, G.nx, G.ev_end, G.twy,
                                                          lds);
        } else {
            row_lines<T, P, GL::L, GL::LSTR, SW_MID, NT>(G.d, it.line0, G.nx,
                                                         G.ph_e[(int)((q - 1) % G.nph)], G.twy,
                                                         lds);
        }
    }
}

template <typename T, int N, bool NT>
__global__ void __launch_bounds__((GroupLaunch<T, N, true>::THREADS))
    k_col_group(const GroupJob *__restrict__ jobs, const GroupItem *__restrict__ items,
                const GroupCounts *__restrict__ counts, const long long *__restrict__ basep,
                int r, int shape)
{
    using GL = GroupLaunch<T, N, true>;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const long long Pidx = *basep + r;
    const GroupItem *list = items + counts->off[4 * shape + 2];
    const int nitems = counts->off[4 * shape + 3];  // a multiple of 8 (padded)
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        // the grid is a multiple of 8, so workgroup b (on XCD b % 8) takes
        // items w = b (mod 8): give each XCD a contiguous range of tiles
        const int we = (w % 8) * (nitems / 8) + w / 8;
        const GroupItem it = list[we];
        if (it.job < 0) continue;
        const GroupJob &G = jobs[it.job];
        const long long q = Pidx - G.first_pass;
        if (q < 0 || q >= (long long)G.nph * G.nsteps) continue;
        col_lines<T, typename GL::Line, GL::L, GL::LSTR, NT, 1>(
            G.d, it.line0, G.ny, G.ph_k[(int)(q % G.nph)], G.twx, lds);
    }
}

// the advance's last node: the next advance starts `passes` passes later
__global__ void __launch_bounds__(64) k_group_tick(long long *basep, int passes)
{
    if (threadIdx.x == 0) *basep += passes;
}

}  // namespace qsim

struct qsim_split_group {
    int prec = 64, method = 2, cap = 0;
    std::vector<qsim_split_job *> jobs;
    std::vector<long long> first, nsteps;
    long long base = 0;  // host copy of the device pass index
    int grid = 0;
    // device tables and their pinned staging
    GroupJob *djobs = nullptr, *hjobs = nullptr;
    GroupItem *dItems = nullptr, *hItems = nullptr;
    size_t item_cap = 0;
    GroupCounts *dcounts = nullptr, *hcounts = nullptr;
    long long *dbase = nullptr;
    bool dirty = true;
    hipEvent_t copied = nullptr;
    bool pending = false;
    // shape classes with member jobs (a capture holds them: a change of the
    // set drops the captures)
    std::vector<int> shapes;
    // grids of each shape's row and column kernels: its item counts when the
    // shape set last changed (multiples of 8, at most 8 workgroups per CU);
    // a shape whose work list grew later loops over it
    int grid_row[kGroupShapes] = {}, grid_col[kGroupShapes] = {};
    size_t shape_points[kGroupShapes] = {};
    // fork / join of the shape chains in direct (uncaptured) advances
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_ev;
    hipEvent_t fork_ev = nullptr;
    // non-temporal line loads of each shape's row and column kernels, chosen
    // with the grids
    bool nt_row[kGroupShapes] = {}, nt_col[kGroupShapes] = {};
    // captured advances, keyed by their launch configuration (pass count and,
    // per shape, grids and load form), oldest first in graph_order; graph_done
    // follows the group's last replay
    std::map<std::vector<long long>, hipGraphExec_t> graphs;
    std::vector<std::vector<long long>> graph_order;
    hipEvent_t graph_done = nullptr;
    bool replayed = false;
    long long captures = 0;
    long long capture_failures = 0;  // advances that ran uncaptured instead
};

namespace {

// A captured advance may still be replaying on the group's stream: wait for
// the last replay (not for the whole device) before destroying captures.
void group_wait_replays(qsim_split_group *g)
{
    if (g->replayed) (void)hipEventSynchronize(g->graph_done);
    g->replayed = false;
}

void group_drop_graphs(qsim_split_group *g)
{
    if (g->graphs.empty()) return;
    group_wait_replays(g);
    for (auto &kv : g->graphs) (void)hipGraphExecDestroy(kv.second);
    g->graphs.clear();
    g->graph_order.clear();
}

// The launch configuration a captured advance of `passes` passes holds.
std::vector<long long> group_graph_key(const qsim_split_group *g, int passes)
{
    std::vector<long long> k{passes};
    for (int sh : g->shapes) {
        k.push_back(sh);
        k.push_back(g->grid_row[sh]);
        k.push_back(g->grid_col[sh]);
        k.push_back(g->nt_row[sh] * 2 + g->nt_col[sh]);
    }
    return k;
}

// lines per workgroup and LDS bytes of one length's group kernels
struct GroupShapeInfo {
    int lrow, lcol;
    size_t lds_row, lds_col;
};

template <typename T>
GroupShapeInfo group_info_t(size_t nx, size_t ny)
{
    GroupShapeInfo r{1, 1, 0, 0};
    switch (ny) {
#define QSIM_GR(NN)                                            \
    case NN:                                                   \
        r.lrow = GroupLaunch<T, NN, false>::L;                 \
        r.lds_row = GroupLaunch<T, NN, false>::LDS_BYTES;      \
        break;
        QSIM_GR(64) QSIM_GR(128) QSIM_GR(256) QSIM_GR(512) QSIM_GR(1024) QSIM_GR(2048)
#undef QSIM_GR
        default: break;
    }
    switch (nx) {
#define QSIM_GC(NN)                                            \
    case NN:                                                   \
        r.lcol = GroupLaunch<T, NN, true>::L;                  \
        r.lds_col = GroupLaunch<T, NN, true>::LDS_BYTES;       \
        break;
        QSIM_GC(64) QSIM_GC(128) QSIM_GC(256) QSIM_GC(512) QSIM_GC(1024) QSIM_GC(2048)
#undef QSIM_GC
        default: break;
    }
    return r;
}

GroupShapeInfo group_info(int prec, size_t nx, size_t ny)
{
    return prec == 64 ? group_info_t<double>(nx, ny) : group_info_t<float>(nx, ny);
}

int group_shape(const qsim_split_job *j)
{
    return group_len_index(j->d.nx) * kGroupLens + group_len_index(j->d.ny);
}

template <typename T, int NX, int NY>
int group_chain_t(qsim_split_group *g, int shape, int passes, hipStream_t s)
{
    using GR = GroupLaunch<T, NY, false>;
    using GC = GroupLaunch<T, NX, true>;
    auto kr = k_row_group<T, NY, false>;
    auto kc = k_col_group<T, NX, false>;
    if constexpr (NY <= 1024)
        if (g->nt_row[shape]) kr = k_row_group<T, NY, true>;
    if constexpr (NX <= 1024)
        if (g->nt_col[shape]) kc = k_col_group<T, NX, true>;
    QSIM_TRY(lds_limit((const void *)kr, GR::LDS_BYTES));
    QSIM_TRY(lds_limit((const void *)kc, GC::LDS_BYTES));
    for (int r = 0; r < passes; ++r) {
        hipLaunchKernelGGL(kr, dim3(g->grid_row[shape]), dim3(GR::THREADS), GR::LDS_BYTES, s,
                           g->djobs, g->dItems, g->dcounts, g->dbase, r, shape);
        QSIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(kc, dim3(g->grid_col[shape]), dim3(GC::THREADS), GC::LDS_BYTES, s,
                           g->djobs, g->dItems, g->dcounts, g->dbase, r, shape);
        QSIM_HIP(hipGetLastError());
    }
    return QSIM_OK;
}

template <typename T, int NX>
int group_chain_nx(qsim_split_group *g, int shape, int passes, hipStream_t s)
{
    switch (shape % kGroupLens) {
        case 0: return group_chain_t<T, NX, 64>(g, shape, passes, s);
        case 1: return group_chain_t<T, NX, 128>(g, shape, passes, s);
        case 2: return group_chain_t<T, NX, 256>(g, shape, passes, s);
        case 3: return group_chain_t<T, NX, 512>(g, shape, passes, s);
        case 4: return group_chain_t<T, NX, 1024>(g, shape, passes, s);
        default: return group_chain_t<T, NX, 2048>(g, shape, passes, s);
    }
}

template <typename T>
int group_chain_prec(qsim_split_group *g, int shape, int passes, hipStream_t s)
{
    switch (shape / kGroupLens) {
        case 0: return group_chain_nx<T, 64>(g, shape, passes, s);
        case 1: return group_chain_nx<T, 128>(g, shape, passes, s);
        case 2: return group_chain_nx<T, 256>(g, shape, passes, s);
        case 3: return group_chain_nx<T, 512>(g, shape, passes, s);
        case 4: return group_chain_nx<T, 1024>(g, shape, passes, s);
        default: return group_chain_nx<T, 2048>(g, shape, passes, s);
    }
}

int group_chain(qsim_split_group *g, int shape, int passes, hipStream_t s)
{
    return g->prec == 64 ? group_chain_prec<double>(g, shape, passes, s)
                         : group_chain_prec<float>(g, shape, passes, s);
}

// One advance on stream s: the shape chains forked onto the group's lanes
// (or, under capture, onto parallel branches of the graph) and joined before
// the pass index moves on.
// One stream and one event per shape chain, created before any capture.
int group_lanes(qsim_split_group *g)
{
    while (g->lanes.size() < g->shapes.size()) {
        hipStream_t ls;
        hipEvent_t le;
        QSIM_HIP(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
        g->lanes.push_back(ls);
        QSIM_HIP(hipEventCreateWithFlags(&le, hipEventDisableTiming));
        g->lane_ev.push_back(le);
    }
    return QSIM_OK;
}

// Replaces the lane streams and events (after a failed capture, which may
// leave the streams it forked to in capture mode).
void group_reset_lanes(qsim_split_group *g)
{
    for (hipStream_t ls : g->lanes) (void)hipStreamDestroy(ls);
    for (hipEvent_t le : g->lane_ev) (void)hipEventDestroy(le);
    g->lanes.clear();
    g->lane_ev.clear();
    (void)hipGetLastError();
}

int group_enqueue(qsim_split_group *g, int passes, hipStream_t s)
{
    const int ns = (int)g->shapes.size();
    QSIM_TRY(group_lanes(g));
    if (ns == 1) {
        QSIM_TRY(group_chain(g, g->shapes[0], passes, s));
    } else if (ns > 1) {
        QSIM_HIP(hipEventRecord(g->fork_ev, s));
        for (int c = 0; c < ns; ++c) {
            QSIM_HIP(hipStreamWaitEvent(g->lanes[c], g->fork_ev, 0));
            QSIM_TRY(group_chain(g, g->shapes[c], passes, g->lanes[c]));
            QSIM_HIP(hipEventRecord(g->lane_ev[c], g->lanes[c]));
            QSIM_HIP(hipStreamWaitEvent(s, g->lane_ev[c], 0));
        }
    }
    hipLaunchKernelGGL(k_group_tick, dim3(1), dim3(64), 0, s, g->dbase, passes);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

// Rebuilds the work lists from the member jobs and uploads them (stream
// ordered before the next advance's kernels).
int group_upload(qsim_split_group *g, hipStream_t s)
{
    const int nj = (int)g->jobs.size();
    std::vector<std::vector<int>> by(kGroupShapes);
    for (int i = 0; i < nj; ++i) by[group_shape(g->jobs[i])].push_back(i);
    std::vector<int> shapes;
    size_t need = 0;
    size_t nrow_of[kGroupShapes] = {}, ncol_of[kGroupShapes] = {};
    for (int sh = 0; sh < kGroupShapes; ++sh) {
        if (by[sh].empty()) continue;
        shapes.push_back(sh);
        size_t nrow = 0, ncol = 0;
        for (int i : by[sh]) {
            const qsim_split_job *j = g->jobs[i];
            const GroupShapeInfo inf = group_info(g->prec, j->d.nx, j->d.ny);
            nrow += (j->d.nx + inf.lrow - 1) / inf.lrow;
            ncol += (j->d.ny + inf.lcol - 1) / inf.lcol;
        }
        nrow_of[sh] = nrow;
        ncol_of[sh] = (ncol + 7) / 8 * 8;
        need += nrow + ncol_of[sh];
    }
    if (shapes != g->shapes) {
        // grids in powers of two (at least 8 workgroups, at most the group's
        // grid), so a shape set that comes back finds its captures again
        g->shapes = shapes;
        auto fit = [g](size_t n) {
            size_t r = 8;
            while (r < n && r < (size_t)g->grid) r *= 2;
            return (int)std::min<size_t>(r, (size_t)g->grid);
        };
        const size_t esz = g->prec == 64 ? sizeof(c128) : sizeof(c64);
        for (int sh : shapes) {
            g->grid_row[sh] = fit(nrow_of[sh]);
            g->grid_col[sh] = fit(ncol_of[sh]);
            g->shape_points[sh] = 0;
            for (int i : by[sh]) g->shape_points[sh] += g->jobs[i]->n;
            // non-temporal line loads by the rule of the standalone sweeps
            // (lines of up to 1024 points, 8 MiB or more per launch)
            const qsim_split_job *j0 = g->jobs[by[sh][0]];
            const size_t bytes = g->shape_points[sh] * esz;
            g->nt_row[sh] = j0->d.ny <= 1024 && line_nt(j0->d.ny, bytes);
            g->nt_col[sh] = j0->d.nx <= 1024 && line_nt(j0->d.nx, bytes);
        }
    }
    if (g->pending) QSIM_HIP(hipEventSynchronize(g->copied));
    g->pending = false;
    if (need > g->item_cap) {
        // new device tables: captured advances hold the old pointer
        group_drop_graphs(g);
        if (g->dItems) QSIM_HIP(hipFree(g->dItems));
        if (g->hItems) QSIM_HIP(hipHostFree(g->hItems));
        g->dItems = nullptr;
        g->hItems = nullptr;
        size_t cap = 1024;
        while (cap < need) cap *= 2;
        QSIM_HIP(hipMalloc(&g->dItems, cap * sizeof(GroupItem)));
        QSIM_HIP(hipHostMalloc(&g->hItems, cap * sizeof(GroupItem), hipHostMallocDefault));
        g->item_cap = cap;
    }
    for (int i = 0; i < nj; ++i) {
        qsim_split_job *j = g->jobs[i];
        GroupJob &G = g->hjobs[i];
        G.d = j->host;
        G.twx = (const c128 *)j->tw_x;
        G.twy = (const c128 *)j->tw_y;
        G.nx = (unsigned)j->d.nx;
        G.ny = (unsigned)j->d.ny;
        G.nph = j->nph;
        G.ev_start = j->ev_start;
        G.ev_end = j->ev_end;
        for (int k = 0; k < 3; ++k) {
            G.ph_k[k] = j->ph_k[k];
            G.ph_e[k] = j->ph_e[k];
        }
        G.first_pass = g->first[i];
        G.nsteps = g->nsteps[i];
    }
    std::memset(g->hcounts, 0, sizeof(GroupCounts));
    size_t w = 0;
    for (int sh : shapes) {
        g->hcounts->off[4 * sh] = (int)w;
        for (int i : by[sh]) {
            const qsim_split_job *j = g->jobs[i];
            const int L = group_info(g->prec, j->d.nx, j->d.ny).lrow;
            for (size_t r0 = 0; r0 < j->d.nx; r0 += L) g->hItems[w++] = GroupItem{i, (unsigned)r0};
        }
        g->hcounts->off[4 * sh + 1] = (int)(w - g->hcounts->off[4 * sh]);
        g->hcounts->off[4 * sh + 2] = (int)w;
        for (int i : by[sh]) {
            const qsim_split_job *j = g->jobs[i];
            const int L = group_info(g->prec, j->d.nx, j->d.ny).lcol;
            for (size_t c0 = 0; c0 < j->d.ny; c0 += L) g->hItems[w++] = GroupItem{i, (unsigned)c0};
        }
        while ((w - g->hcounts->off[4 * sh + 2]) % 8) g->hItems[w++] = GroupItem{-1, 0};
        g->hcounts->off[4 * sh + 3] = (int)(w - g->hcounts->off[4 * sh + 2]);
    }
    if (nj)
        QSIM_HIP(hipMemcpyAsync(g->djobs, g->hjobs, sizeof(GroupJob) * nj, hipMemcpyHostToDevice,
                                s));
    if (w)
        QSIM_HIP(hipMemcpyAsync(g->dItems, g->hItems, sizeof(GroupItem) * w,
                                hipMemcpyHostToDevice, s));
    QSIM_HIP(hipMemcpyAsync(g->dcounts, g->hcounts, sizeof(GroupCounts), hipMemcpyHostToDevice,
                            s));
    QSIM_HIP(hipEventRecord(g->copied, s));
    g->pending = true;
    g->dirty = false;
    return QSIM_OK;
}

int group_find(const qsim_split_group *g, const qsim_split_job *j)
{
    for (size_t i = 0; i < g->jobs.size(); ++i)
        if (g->jobs[i] == j) return (int)i;
    return -1;
}

}  // namespace

extern "C" {

int qsim_hip_split_group_destroy(qsim_split_group *g)
{
    QSIM_DEVICE_WAIT_GUARD;
    if (!g) return QSIM_OK;
    if (g->pending) (void)hipEventSynchronize(g->copied);
    group_drop_graphs(g);
    (void)hipDeviceSynchronize();
    for (hipStream_t ls : g->lanes) (void)hipStreamDestroy(ls);
    for (hipEvent_t le : g->lane_ev) (void)hipEventDestroy(le);
    if (g->fork_ev) (void)hipEventDestroy(g->fork_ev);
    if (g->graph_done) (void)hipEventDestroy(g->graph_done);
    if (g->copied) (void)hipEventDestroy(g->copied);
    if (g->djobs) (void)hipFree(g->djobs);
    if (g->hjobs) (void)hipHostFree(g->hjobs);
    if (g->dItems) (void)hipFree(g->dItems);
    if (g->hItems) (void)hipHostFree(g->hItems);
    if (g->dcounts) (void)hipFree(g->dcounts);
    if (g->hcounts) (void)hipHostFree(g->hcounts);
    if (g->dbase) (void)hipFree(g->dbase);
    delete g;
    return QSIM_OK;
}

int qsim_hip_split_group_create(qsim_split_group **out, int capacity, int precision, int method)
{
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(out && capacity > 0, "group: null out-pointer or capacity < 1");
    QSIM_CHECK(precision == 64 || precision == 32, "group: precision must be 64 or 32");
    QSIM_CHECK(method == 2 || method == 4, "group: method must be 2 (O2) or 4 (O4)");
    *out = nullptr;
    qsim_split_group *g = new qsim_split_group();
    g->prec = precision;
    g->method = method;
    g->cap = capacity;
    const int cu = device_cu_count();
    // the largest grid of a shape chain's kernels: eight workgroups per CU
    g->grid = ((cu > 0 ? 8 * cu : 2048) + 7) / 8 * 8;
    int rc = QSIM_OK;
    const long long zero = 0;
    if (hipMalloc(&g->djobs, sizeof(GroupJob) * capacity) != hipSuccess ||
        hipHostMalloc(&g->hjobs, sizeof(GroupJob) * capacity, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&g->dcounts, sizeof(GroupCounts)) != hipSuccess ||
        hipHostMalloc(&g->hcounts, sizeof(GroupCounts), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&g->dbase, sizeof(long long)) != hipSuccess ||
        hipEventCreateWithFlags(&g->copied, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->graph_done, hipEventDisableTiming) != hipSuccess ||
        hipMemcpy(g->dbase, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        rc = set_error(QSIM_ERR_NOMEM, "group: allocation failed");
    }
    if (rc != QSIM_OK) {
        qsim_hip_split_group_destroy(g);
        return rc;
    }
    *out = g;
    return QSIM_OK;
}

int qsim_hip_split_group_add(qsim_split_group *g, qsim_split_job *j, size_t nsteps)
{
    QSIM_CHECK(g && j, "group: null argument");
    QSIM_CHECK(nsteps >= 1, "group: a job needs at least one step");
    QSIM_CHECK((int)g->jobs.size() < g->cap, "group: full");
    QSIM_CHECK(group_find(g, j) < 0, "group: the job is already in the group");
    QSIM_CHECK(j->fused && j->d.dim == 2 && j->d.precision == g->prec && j->d.method == g->method,
               "group: needs a fused 2D job of the group's precision and method");
    QSIM_CHECK(group_len_ok(j->d.nx) && group_len_ok(j->d.ny),
               "group: grid axes must be powers of two from 64 to 2048");
    QSIM_CHECK(j->xsplit == 1 && j->row_var == default_variant(g->prec, j->d.ny, true) &&
                   j->col_var == default_variant(g->prec, j->d.nx, false),
               "group: the job must keep its default launch schedule");
    g->jobs.push_back(j);
    g->first.push_back(g->base);
    g->nsteps.push_back((long long)nsteps);
    g->dirty = true;
    return QSIM_OK;
}

int qsim_hip_split_group_remove(qsim_split_group *g, qsim_split_job *j)
{
    QSIM_CHECK(g && j, "group: null argument");
    const int i = group_find(g, j);
    QSIM_CHECK(i >= 0, "group: the job is not in the group");
    g->jobs.erase(g->jobs.begin() + i);
    g->first.erase(g->first.begin() + i);
    g->nsteps.erase(g->nsteps.begin() + i);
    g->dirty = true;
    return QSIM_OK;
}

int qsim_hip_split_group_advance(qsim_split_group *g, int passes, int use_graph,
                                 qsim_stream_t stream)
{
    qsim::TraceRange tr_("qsim_hip_split_group_advance");
    QSIM_CHECK(g && passes >= 1, "group: null group or passes < 1");
    hipStream_t s = as_stream(stream);
    // uploads (re)allocate the work lists, captures must not meet another
    // thread's device-wide wait
    std::unique_lock<std::recursive_mutex> lk(device_wait_lock(), std::defer_lock);
    if (g->dirty) {
        lk.lock();
        QSIM_TRY(group_upload(g, s));
    }
    if (!use_graph) {
        QSIM_TRY(group_enqueue(g, passes, s));
    } else {
        const std::vector<long long> key = group_graph_key(g, passes);
        auto it = g->graphs.find(key);
        if (it == g->graphs.end()) {
            if (!lk.owns_lock()) lk.lock();
            if (g->graphs.size() >= kMaxGraphs) {
                group_wait_replays(g);
                auto old = g->graphs.find(g->graph_order.front());
                (void)hipGraphExecDestroy(old->second);
                g->graphs.erase(old);
                g->graph_order.erase(g->graph_order.begin());
            }
            QSIM_TRY(group_lanes(g));
            hipStream_t cs;
            QSIM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            int rc = QSIM_OK;
            if (e == hipSuccess) {
                rc = group_enqueue(g, passes, cs);
                hipError_t e2 = hipStreamEndCapture(cs, &gr);
                if (rc == QSIM_OK && e2 != hipSuccess) rc = hip_error(e2, "hipStreamEndCapture");
            } else {
                rc = hip_error(e, "hipStreamBeginCapture");
            }
            hipGraphExec_t ex = nullptr;
            if (rc == QSIM_OK) {
                e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                if (e != hipSuccess) rc = hip_error(e, "hipGraphInstantiate");
            }
            if (gr) (void)hipGraphDestroy(gr);
            (void)hipStreamDestroy(cs);
            if (rc != QSIM_OK) {
                // the runtime refused or invalidated the capture: this advance
                // runs as direct launches (the same kernels, bitwise the same
                // results) on fresh lanes, and the next one captures again
                group_reset_lanes(g);
                ++g->capture_failures;
                QSIM_TRY(group_enqueue(g, passes, s));
                g->base += passes;
                return QSIM_OK;
            }
            it = g->graphs.emplace(key, ex).first;
            g->graph_order.push_back(key);
            ++g->captures;
        }
        QSIM_HIP(hipGraphLaunch(it->second, s));
        QSIM_HIP(hipEventRecord(g->graph_done, s));
        g->replayed = true;
    }
    g->base += passes;
    return QSIM_OK;
}

int qsim_hip_split_group_info(const qsim_split_group *g, int *jobs, int *shapes, int *graphs,
                              long long *capture_failures)
{
    QSIM_CHECK(g, "group: null argument");
    if (jobs) *jobs = (int)g->jobs.size();
    if (shapes) *shapes = (int)g->shapes.size();
    if (graphs) *graphs = (int)g->graphs.size();
    if (capture_failures) *capture_failures = g->capture_failures;
    return QSIM_OK;
}

int qsim_hip_split_group_status(const qsim_split_group *g, const qsim_split_job *j,
                                int *finished, long long *passes_left)
{
    QSIM_CHECK(g && j, "group: null argument");
    const int i = group_find(g, j);
    QSIM_CHECK(i >= 0, "group: the job is not in the group");
    const long long last = g->first[i] + (long long)j->nph * g->nsteps[i];
    if (finished) *finished = last < g->base;
    if (passes_left) *passes_left = last < g->base ? 0 : last + 1 - g->base;
    return QSIM_OK;
}

}  // extern "C"
