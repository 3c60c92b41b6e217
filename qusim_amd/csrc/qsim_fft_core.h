// qsim_fft_core.h -- FFT building blocks for gfx950.
//
// * dft_reg<T, R, INV>: an R-point DFT (R = 2..64, power of two) held in one
//   thread's registers: radix-2 decimation in frequency with compile-time
//   twiddles, output permuted back to natural order by register renaming.
// * Pow2Line<T, N, RA, RB, RC, RD>: one length-N line transformed by N/RMAX
//   threads in up to four Stockham passes (radices RA, RB, RC, RD), data in
//   registers, one LDS exchange between passes.  Forward runs RA..RD;
//   inverse runs RD..RA, so the forward's outputs are exactly the
//   inverse's inputs in every thread's registers (and vice versa): an
//   FFT -> pointwise -> IFFT sequence needs no exchange around the pointwise
//   step.
//
// Sign convention: forward X_k = sum x_n e^{-2 pi i k n / N}, inverse +,
// unnormalised (reference src/FourierTransform.h:16-19).
#pragma once

#include "qsim_common.h"
#include "twiddle_consts.h"

namespace qsim {

template <typename T>
struct TwC;
template <>
struct TwC<double> {
    static __device__ __forceinline__ double c(int k) { return kCos64d[k & 63]; }
    static __device__ __forceinline__ double s(int k) { return kSin64d[k & 63]; }
};
template <>
struct TwC<float> {
    static __device__ __forceinline__ float c(int k) { return kCos64f[k & 63]; }
    static __device__ __forceinline__ float s(int k) { return kSin64f[k & 63]; }
    static __device__ __forceinline__ float cl(int k) { return kCos64fl[k & 63]; }
    static __device__ __forceinline__ float sl(int k) { return kSin64fl[k & 63]; }
};

// FP32 data times a float-float factor (hi + lo): the factor carries ~48
// bits, so a rounded constant never repeats its error coherently step after
// step; only the products round (unbiased).
__device__ __forceinline__ c64 mul_ff(c64 a, c64 h, c64 l)
{
    const float sr = a.x * l.x - a.y * l.y, si = a.x * l.y + a.y * l.x;
    return mk(fmaf(a.x, h.x, fmaf(-a.y, h.y, sr)), fmaf(a.x, h.y, fmaf(a.y, h.x, si)));
}

// a * exp(sgn * 2 pi i * k / m), sgn = +1 for INV, -1 forward; k, m are
// compile-time after unrolling, so every branch folds away.
template <typename T, bool INV>
__device__ __forceinline__ cx<T> twmul(cx<T> a, int k, int m)
{
    const int q = (k * (64 / m)) & 63;  // angle 2 pi q / 64
    if (q == 0) return a;
    if (q == 32) return mk(-a.x, -a.y);
    if (q == 16) return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);   // * (+-i)
    if (q == 48) return INV ? mk(a.y, -a.x) : mk(-a.y, a.x);
    const T c = TwC<T>::c(q);
    const T s = INV ? TwC<T>::s(q) : -TwC<T>::s(q);
    if constexpr (sizeof(T) == 4) {
        const float cl = TwC<float>::cl(q);
        const float sl = INV ? TwC<float>::sl(q) : -TwC<float>::sl(q);
        if ((q & 15) == 8) {  // |c| = |s| = sqrt(1/2): (hi + lo) * (a.x -+ a.y)
            const float h = c < 0 ? -c : c, hl = c < 0 ? -cl : cl;
            const float pc = c < 0 ? -a.x : a.x, qc = c < 0 ? -a.y : a.y;
            const float ps = s < 0 ? -a.y : a.y, qs = s < 0 ? -a.x : a.x;
            const float d1 = pc - ps, d2 = qc + qs;
            return mk(fmaf(h, d1, hl * d1), fmaf(h, d2, hl * d2));
        }
        return mul_ff(a, mk(c, s), mk(cl, sl));
    } else {
        if ((q & 15) == 8) {  // |c| = |s| = sqrt(1/2)
            const T h = c < 0 ? -c : c;
            const T pc = c < 0 ? -a.x : a.x, qc = c < 0 ? -a.y : a.y;   // c*a / h
            const T ps = s < 0 ? -a.y : a.y, qs = s < 0 ? -a.x : a.x;   // s*a / h
            return mk(h * (pc - ps), h * (qc + qs));
        }
        return mk(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
__host__ __device__ constexpr int brev(int x, int bits)
{
    int r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// In-register R-point DFT, natural order in and out.
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_reg(cx<T> *v)
{
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (R == 4) {
        cx<T> a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = v[1] - v[3];
        // forward: X1 = b - i d, X3 = b + i d
        cx<T> id = INV ? mk(-d.y, d.x) : mk(d.y, -d.x);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + id;
        v[3] = b - id;
    } else {
#pragma unroll
        for (int m = R; m >= 2; m >>= 1) {
            const int h = m >> 1;
#pragma unroll
            for (int b = 0; b < R; b += m) {
#pragma unroll
                for (int k = 0; k < h; ++k) {
                    cx<T> a = v[b + k], c = v[b + k + h];
                    v[b + k] = a + c;
                    v[b + k + h] = twmul<T, INV>(a - c, k, m);
                }
            }
        }
        cx<T> t[R];
        constexpr int bits = ilog2(R);
#pragma unroll
        for (int r = 0; r < R; ++r) t[r] = v[brev(r, bits)];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = t[r];
    }
}

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// One length-N line, N = RA*RB*RC*RD (RB, RC, RD may be 1; RD > 1 needs
// RC > 1).  Thread t < TPL holds E = RMAX elements; for a pass of radix R it
// owns butterflies j = t + u*TPL, u < E/R, element r of butterfly u in
// v[u*R + r].  More passes mean smaller E: more threads (waves) per line.
// TREE: build each butterfly's twiddle powers as a product tree (w^r =
// w^(r/2) w^(r - r/2), all R-1 powers live) instead of the recurrence
// w^r = w^(r-1) w (one live power).  Same multiplies, other register
// allocation; a kernel opts in where that avoids spills.
// SWZ: LDS placement of the exchanges.  0 pads one slot per R elements
// after a radix-R pass; 1, 2 and 3 XOR-swizzle the low bits of the element
// index inside its aligned 8- or 16-element block instead (no padding):
// m ^ ((m >> 3) & 7), m ^ ((m >> 2) & 7), m ^ ((m >> 3) & 15); 4 XORs the
// low 3 bits with a per-line key (0..7) that the caller passes (the line's
// index in a line-fast tile).  Padding keeps the stores free of conflicts
// but breaks the runs of consecutive addresses that the loads want; a
// swizzle can keep both.
template <typename T, int N, int RA, int RB, int RC, int RD = 1, bool TREE = false, int SWZ = 0>
struct Pow2Line {
    static_assert(RA * RB * RC * RD == N, "radices must multiply to N");
    static_assert(RD == 1 || RC > 1, "a fourth pass needs a third");
    static constexpr int NP = RD > 1 ? 4 : (RC > 1 ? 3 : (RB > 1 ? 2 : 1));
    static constexpr int RA_ = RA, RB_ = RB, RC_ = RC, RD_ = RD;
    // the same line with another LDS placement
    template <int S>
    using WithSwz = Pow2Line<T, N, RA, RB, RC, RD, TREE, S>;
    static constexpr int E = cmax(cmax(RA, RB), cmax(RC, RD));
    static constexpr int TPL = N / E;
    static constexpr int RMIN = cmin(cmin(RA, RB > 1 ? RB : RA),
                                     cmin(RC > 1 ? RC : RA, RD > 1 ? RD : RA));
    // LDS elements per line: N (swizzled) or N plus one pad per RMIN, rounded
    // to 1 mod 16 so lines placed side by side start on different banks.
    static constexpr int LS_RAW = SWZ == 0 ? N + N / RMIN + 1 : N;
    static_assert(SWZ != 4 || N % 8 == 0, "per-line keys permute aligned 8-element blocks");
    static constexpr int LS = LS_RAW + ((1 - LS_RAW % 16) + 16) % 16;

    // LDS index of line element m in the exchange after a radix-R pass;
    // the swizzles stay inside m's aligned block, so indices stay below N
    template <int R>
    static __host__ __device__ __forceinline__ int pad(int m, int key)
    {
        if constexpr (SWZ == 0) return m + m / R;
        else if constexpr (SWZ == 1) return m ^ ((m >> 3) & 7);
        else if constexpr (SWZ == 2) return m ^ ((m >> 2) & 7);
        else if constexpr (SWZ == 3) return m ^ ((m >> 3) & 15);
        else return m ^ (key & 7);
    }

    // Positions of the elements thread t holds as inputs of a first pass of
    // radix R: element (u, r) is at j + r*(N/R).
    template <int R>
    static __device__ __forceinline__ int in_pos(int t, int u, int r)
    {
        return t + u * TPL + r * (N / R);
    }

    // twiddle + butterfly for a pass of radix R after NS points combined:
    // one FP64 table load per butterfly, the powers w^r by FP64
    // multiplication (error ~r ulp, far below the 1e-12 bar).  FP32 data is
    // widened, multiplied in FP64 and rounded once: no rounded twiddle is
    // applied again at every step, and it takes half the instructions of a
    // float-float product.
    template <int R, int NS, bool INV>
    static __device__ __forceinline__ void pass(cx<T> *v, int t, const c128 *__restrict__ tw)
    {
#pragma unroll
        for (int u = 0; u < E / R; ++u) {
            if constexpr (NS > 1) {
                const int j = t + u * TPL;
                const int k = j & (NS - 1);
                const int step = k * (N / (NS * R));
                if constexpr (TREE) {
                    c128 w[R];
                    w[1] = tw[step];
#pragma unroll
                    for (int r = 2; r < R; ++r) w[r] = w[r / 2] * w[r - r / 2];
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        if constexpr (sizeof(T) == 8) {
                            v[u * R + r] = INV ? mulconj(v[u * R + r], w[r]) : v[u * R + r] * w[r];
                        } else {
                            const c128 x = cvt<double>(v[u * R + r]);
                            v[u * R + r] = cvt<float>(INV ? mulconj(x, w[r]) : x * w[r]);
                        }
                    }
                } else if constexpr (sizeof(T) == 8) {
                    const cx<T> w1 = tw[step];
                    cx<T> w = w1;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        v[u * R + r] = INV ? mulconj(v[u * R + r], w) : v[u * R + r] * w;
                        if (r + 1 < R) w = w * w1;
                    }
                } else {
                    // the same FP64 powers; the product rounds once
                    const c128 w1 = tw[step];
                    c128 w = w1;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const c128 x = cvt<double>(v[u * R + r]);
                        v[u * R + r] = cvt<float>(INV ? mulconj(x, w) : x * w);
                        if (r + 1 < R) w = w * w1;
                    }
                }
            }
            dft_reg<T, R, INV>(v + u * R);
        }
    }

    // LDS exchange from the outputs of a pass (radix R, after NS) to the
    // inputs of a pass of radix R2.
    template <int R, int NS, int R2>
    static __device__ __forceinline__ void exchange(cx<T> *v, int t, cx<T> *lds, int key)
    {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < E / R; ++u) {
            const int j = t + u * TPL;
            const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
            for (int r = 0; r < R; ++r) lds[pad<R>(base + r * NS, key)] = v[u * R + r];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < E / R2; ++u)
#pragma unroll
            for (int r = 0; r < R2; ++r) v[u * R2 + r] = lds[pad<R>(in_pos<R2>(t, u, r), key)];
    }

    // Forward transform.  In: inputs of radix-RA pass (in_pos<RA>).  Out: X at
    // in_pos<RLAST>(t, u, r) with RLAST = RD, RC, RB or RA (NP = 4, 3, 2, 1).
    static __device__ __forceinline__ void forward(cx<T> *v, int t, cx<T> *lds,
                                                   const c128 *__restrict__ tw, int key = 0)
    {
        pass<RA, 1, false>(v, t, tw);
        if constexpr (NP >= 2) {
            exchange<RA, 1, RB>(v, t, lds, key);
            pass<RB, RA, false>(v, t, tw);
        }
        if constexpr (NP >= 3) {
            exchange<RB, RA, RC>(v, t, lds, key);
            pass<RC, RA * RB, false>(v, t, tw);
        }
        if constexpr (NP == 4) {
            exchange<RC, RA * RB, RD>(v, t, lds, key);
            pass<RD, RA * RB * RC, false>(v, t, tw);
        }
    }

    static constexpr int RLAST = NP == 4 ? RD : (NP == 3 ? RC : (NP == 2 ? RB : RA));

    // Inverse transform, radices RLAST, ..., RA.  In: in_pos<RLAST>.  Out: x at
    // in_pos<RA>(t, u, r).
    static __device__ __forceinline__ void inverse(cx<T> *v, int t, cx<T> *lds,
                                                   const c128 *__restrict__ tw, int key = 0)
    {
        if constexpr (NP == 1) {
            pass<RA, 1, true>(v, t, tw);
        } else if constexpr (NP == 2) {
            pass<RB, 1, true>(v, t, tw);
            exchange<RB, 1, RA>(v, t, lds, key);
            pass<RA, RB, true>(v, t, tw);
        } else if constexpr (NP == 3) {
            pass<RC, 1, true>(v, t, tw);
            exchange<RC, 1, RB>(v, t, lds, key);
            pass<RB, RC, true>(v, t, tw);
            exchange<RB, RC, RA>(v, t, lds, key);
            pass<RA, RC * RB, true>(v, t, tw);
        } else {
            pass<RD, 1, true>(v, t, tw);
            exchange<RD, 1, RC>(v, t, lds, key);
            pass<RC, RD, true>(v, t, tw);
            exchange<RC, RD, RB>(v, t, lds, key);
            pass<RB, RD * RC, true>(v, t, tw);
            exchange<RB, RD * RC, RA>(v, t, lds, key);
            pass<RA, RD * RC * RB, true>(v, t, tw);
        }
    }
};

// ---------------------------------------------------------------------------
// Line addressing: a tensor viewed as [outer][N][S]; line g (g < outer*S) is
// (o = g / S, c = g % S), element m at o*N*S + c + m*S.
__device__ __forceinline__ size_t line_base(size_t g, size_t N, size_t S)
{
    return (g / S) * N * S + (g % S);
}

}  // namespace qsim
