// qsim_fft_kernels.h -- FFT kernels over the lines of one tensor axis.
#pragma once

#include "qsim_fft_core.h"

namespace qsim {

// Radix plan of the register-resident power-of-two line FFT, per size,
// element type and launch variant V (variant 0 is the default; the others
// are the candidates the split-step autotuner times).  N/E threads per line,
// E = largest radix.  ROW_L / COL_L: lines per workgroup of the row
// (contiguous) and column (strided) kernels; 0 = derive from LDS.
template <typename T, int N, int V = 0>
struct P2Cfg;
#define QSIM_P2CFG4(TYPE, N_, V_, A, B, C_, D_, RL, CL) \
    template <>                                          \
    struct P2Cfg<TYPE, N_, V_> {                         \
        using Line = Pow2Line<TYPE, N_, A, B, C_, D_>;   \
        static constexpr int ROW_L = RL, COL_L = CL;     \
    };
#define QSIM_P2CFG(TYPE, N_, V_, A, B, C_, RL, CL) QSIM_P2CFG4(TYPE, N_, V_, A, B, C_, 1, RL, CL)
QSIM_P2CFG(double, 16, 0, 4, 4, 1, 0, 0)
QSIM_P2CFG(double, 32, 0, 8, 4, 1, 0, 0)
QSIM_P2CFG(double, 64, 0, 8, 8, 1, 0, 0)
QSIM_P2CFG(double, 128, 0, 16, 8, 1, 0, 0)
QSIM_P2CFG(double, 128, 1, 8, 8, 2, 0, 0)  // 8 per thread (Born sweeps)
QSIM_P2CFG(double, 256, 0, 16, 16, 1, 0, 0)
QSIM_P2CFG(double, 256, 1, 8, 8, 4, 0, 0)  // 8 per thread (Born sweeps)
QSIM_P2CFG(double, 512, 0, 8, 8, 8, 0, 0)
QSIM_P2CFG(double, 512, 1, 8, 8, 8, 0, 0)  // Born sweeps (as variant 0)
QSIM_P2CFG(double, 1024, 0, 16, 16, 4, 1, 4)
QSIM_P2CFG(double, 1024, 1, 16, 16, 4, 4, 4)
QSIM_P2CFG(double, 1024, 2, 32, 32, 1, 2, 4)
QSIM_P2CFG(double, 1024, 3, 32, 32, 1, 2, 8)
QSIM_P2CFG(double, 1024, 4, 8, 8, 16, 1, 4)
QSIM_P2CFG(double, 1024, 5, 16, 16, 4, 2, 8)
// four passes, 8 elements per thread: two waves per line
QSIM_P2CFG4(double, 1024, 6, 8, 8, 4, 4, 1, 4)
QSIM_P2CFG4(double, 1024, 7, 8, 8, 4, 4, 2, 8)
QSIM_P2CFG(double, 2048, 0, 16, 16, 8, 1, 4)
QSIM_P2CFG(double, 2048, 1, 16, 16, 8, 2, 2)
QSIM_P2CFG(double, 2048, 2, 32, 32, 2, 1, 4)
QSIM_P2CFG(double, 2048, 3, 8, 16, 16, 1, 4)
QSIM_P2CFG(double, 2048, 4, 16, 16, 8, 1, 8)
QSIM_P2CFG(double, 2048, 5, 16, 16, 8, 4, 4)
QSIM_P2CFG4(double, 2048, 6, 8, 8, 8, 4, 1, 4)
QSIM_P2CFG4(double, 2048, 7, 8, 8, 8, 4, 1, 2)
QSIM_P2CFG(double, 4096, 0, 16, 16, 16, 0, 0)
QSIM_P2CFG4(double, 4096, 1, 8, 8, 8, 8, 0, 0)  // 1D split jobs: 8 per thread
// 16 elements per thread keeps these within 128 registers (32 per thread
// spilled ~1 KiB per lane); radix 8 passes keep the padding within 160 KiB
QSIM_P2CFG4(double, 8192, 0, 16, 8, 8, 8, 0, 0)  // 1D split jobs only
QSIM_P2CFG(float, 16, 0, 4, 4, 1, 0, 0)
QSIM_P2CFG(float, 32, 0, 8, 4, 1, 0, 0)
QSIM_P2CFG(float, 64, 0, 8, 8, 1, 0, 0)
QSIM_P2CFG(float, 128, 0, 16, 8, 1, 0, 0)
QSIM_P2CFG(float, 128, 1, 8, 8, 2, 0, 0)  // 8 per thread (Born sweeps)
QSIM_P2CFG(float, 256, 0, 16, 16, 1, 0, 0)
QSIM_P2CFG(float, 256, 1, 8, 8, 4, 0, 0)  // 8 per thread (Born sweeps)
QSIM_P2CFG(float, 512, 0, 16, 32, 1, 0, 0)
// 8 per thread (Born sweeps); 8-line column tiles (512 threads): with 16
// (1024 threads at 128 registers) the tile kept 20 B per lane in scratch, and
// 512^2 FP32 took 16.1 against 12.8 us per step (2D Born 20.0 against 16.8 us
// per iteration)
QSIM_P2CFG(float, 512, 1, 8, 8, 8, 0, 8)
QSIM_P2CFG(float, 1024, 0, 32, 32, 1, 2, 8)
QSIM_P2CFG(float, 1024, 1, 32, 32, 1, 4, 16)
QSIM_P2CFG(float, 1024, 2, 16, 16, 4, 1, 8)
QSIM_P2CFG(float, 1024, 3, 16, 16, 4, 2, 16)
QSIM_P2CFG(float, 1024, 4, 32, 32, 1, 2, 4)
QSIM_P2CFG(float, 1024, 5, 16, 16, 4, 4, 4)
QSIM_P2CFG4(float, 1024, 6, 8, 8, 4, 4, 2, 8)
QSIM_P2CFG4(float, 1024, 7, 8, 8, 4, 4, 1, 4)
QSIM_P2CFG(float, 2048, 0, 16, 16, 8, 1, 8)
QSIM_P2CFG(float, 2048, 1, 32, 64, 1, 1, 8)
QSIM_P2CFG(float, 2048, 2, 64, 32, 1, 1, 8)
QSIM_P2CFG(float, 2048, 3, 16, 16, 8, 2, 16)
QSIM_P2CFG(float, 2048, 4, 32, 32, 2, 1, 8)
QSIM_P2CFG(float, 2048, 5, 16, 16, 8, 1, 4)
QSIM_P2CFG4(float, 2048, 6, 8, 8, 8, 4, 1, 4)
QSIM_P2CFG4(float, 2048, 7, 8, 8, 8, 4, 1, 2)
QSIM_P2CFG(float, 4096, 0, 16, 16, 16, 0, 0)
QSIM_P2CFG4(float, 4096, 1, 8, 8, 8, 8, 0, 0)  // 1D split jobs: 8 per thread
QSIM_P2CFG4(float, 8192, 0, 16, 8, 8, 8, 0, 0)
QSIM_P2CFG(float, 16384, 0, 16, 32, 32, 0, 0)  // 1D split jobs only (32 per thread: 512 threads, 256 registers)
#undef QSIM_P2CFG
#undef QSIM_P2CFG4

__host__ __device__ constexpr int pow2floor(int x) { return x <= 1 ? 1 : 2 * pow2floor(x / 2); }

// Lines per workgroup: rows (position-fast) fill ~256 threads; strided axes
// (line-fast) take 128 B of adjacent lines per row of the tile; both kept
// within 96 KiB of LDS and a power of two, unless the variant fixes them.
template <typename T, int N, bool LF, int V = 0>
struct P2Launch {
    using Cfg = P2Cfg<T, N, V>;
    using Base = typename Cfg::Line;
    // LDS placement of the exchanges (Pow2Line SWZ), each measured per store
    // and load pattern with rocprofv3 (SQ_LDS_IDX_ACTIVE per instruction):
    // rows with a wave or more per line take an XOR swizzle (FP64 1024-point
    // rows 1.33-1.61 -> 1.11-1.17 times the conflict-free LDS cycles, FP32
    // 1.56 -> 1.22); FP64 line-fast tiles of 8 lines XOR each line's element
    // index with the line's index and drop the padding (1.33 -> 1.00 for
    // every 8-line plan measured); other tiles and shorter rows keep the
    // padding.
    static constexpr int SWZ_ROW = (!LF && Base::TPL >= 64) ? (sizeof(cx<T>) == 16 ? 1 : 3) : 0;
    using Sized = typename Base::template WithSwz<SWZ_ROW>;
    static constexpr int LF_LINES = 128 / (int)sizeof(cx<T>);
    static constexpr int ROWS = cmax(1, 256 / Sized::TPL);
    static constexpr int L_WANT = LF ? LF_LINES : ROWS;
    static constexpr int LS_MAX = Sized::LS_RAW + 15;
    static constexpr int L_LDS = cmax(1, (96 * 1024) / (LS_MAX * (int)sizeof(cx<T>)));
    static constexpr int L_FIXED = LF ? Cfg::COL_L : Cfg::ROW_L;
    // never more lines than the 160 KiB of LDS one workgroup can hold
    static constexpr int L_CAP = pow2floor(cmax(1, (160 * 1024) / (LS_MAX * (int)sizeof(cx<T>))));
    static constexpr int L = cmin(L_CAP, L_FIXED > 0 ? L_FIXED : pow2floor(cmin(L_WANT, L_LDS)));
    static constexpr int SWZ = (LF && sizeof(cx<T>) == 16 && L == 8) ? 4 : SWZ_ROW;
    using Line = typename Base::template WithSwz<SWZ>;
    static constexpr int THREADS = L * Line::TPL;
    // LDS stride between the lines of a workgroup (in elements).  Line-fast
    // tiles put lane i on line i % L, element i / L.  The stride (mod 16)
    // is the one that took the fewest LDS cycles over every exchange of
    // these tiles, measured per element size with rocprofv3
    // (SQ_LDS_IDX_ACTIVE per instruction, stores and loads timed apart):
    // 8-byte elements 16/L; 16-byte elements 8/L for L = 2 and 4 (4-line
    // tiles 1.67 -> 1.33 times the conflict-free cycles) and 8 with the
    // per-line key for L = 8.  Rows keep 1 (mod 16).
    static constexpr int TARGET =
        !LF ? 1
            : (sizeof(cx<T>) == 8 ? (L >= 16 ? 1 : 16 / L)
                                  : (SWZ == 4 ? 8 : (L >= 8 ? 1 : 8 / L)));
    static constexpr int LSTR =
        Line::LS_RAW + (((TARGET - Line::LS_RAW % 16) % 16) + 16) % 16;
    static constexpr size_t LDS_BYTES = (size_t)L * LSTR * sizeof(cx<T>);
};

// Radix plan of the Born sweeps per length: 8 elements per thread from 128
// to 4096 points, so a solve of one energy on a few hundred lines still puts
// several waves on every SIMD.
template <int N>
struct LineV {
    static constexpr int V = (N == 1024 || N == 2048) ? 6
                             : (N == 128 || N == 256 || N == 512 || N == 4096) ? 1 : 0;
};

// Radix plan of the FFT plan kernels: the default plans, except FP32 512 and
// 1024, whose 16-32-element plans leave a 2D transform of that size a
// fraction of a wave per SIMD (measured: 512^2 15.0 -> 10.5 us, 1024^2 23.0
// -> 20.4 us; at 2048 and for FP64 the default plans are faster).
template <typename T, int N>
struct FftV {
    static constexpr int V = sizeof(T) == 4 && (N == 512 || N == 1024) ? LineV<N>::V : 0;
};

// Plain transform of `nlines` lines of length N with stride S (in may be out).
// NT: non-temporal loads and stores (global address space), which the host
// picks for passes of 8-16 MiB: there they keep the twiddles and the other
// lines' data in L2 (1024^2 FP64 2D transform 27.4 -> 21.2 us, 128^3 FP32
// 27.3 -> 22 us); smaller passes stay in L2 anyway, and at 32-64 MiB they
// cost up to 14%, so those keep the plain accesses.
template <typename T, int N, bool LF, bool INV, bool NT = false>
__global__ void __launch_bounds__((P2Launch<T, N, LF, FftV<T, N>::V>::THREADS))
    k_fft_p2(const cx<T> *in, cx<T> *out, size_t S, size_t nlines,
             const c128 *__restrict__ tw, T scale)
{
    using LP = P2Launch<T, N, LF, FftV<T, N>::V>;
    using P = typename LP::Line;
    constexpr int L = LP::L;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int tid = threadIdx.x;
    const int l = LF ? tid % L : tid / P::TPL;
    const int t = LF ? tid / L : tid % P::TPL;
    const size_t g = (size_t)blockIdx.x * L + l;
    const bool active = g < nlines;
    const size_t base = active ? line_base(g, N, S) : 0;
    constexpr int R0 = INV ? P::RLAST : P::RA_;
    constexpr int R1 = INV ? P::RA_ : P::RLAST;
    cx<T> v[P::E];
#pragma unroll
    for (int u = 0; u < P::E / R0; ++u)
#pragma unroll
        for (int r = 0; r < R0; ++r)
            v[u * R0 + r] =
                !active ? mk<T>(0, 0)
                        : NT ? ld_nt(in + base + (size_t)P::template in_pos<R0>(t, u, r) * S)
                             : in[base + (size_t)P::template in_pos<R0>(t, u, r) * S];
    if constexpr (INV)
        P::inverse(v, t, lds + l * LP::LSTR, tw, l);
    else
        P::forward(v, t, lds + l * LP::LSTR, tw, l);
    if (active) {
#pragma unroll
        for (int u = 0; u < P::E / R1; ++u)
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                cx<T> *o = &out[base + (size_t)P::template in_pos<R1>(t, u, r) * S];
                if constexpr (NT) store_nt(o, v[u * R1 + r] * scale);
                else *o = v[u * R1 + r] * scale;
            }
    }
}

// Four-step (Bailey) passes of one long axis of length n = n1 * n2 over a
// tensor [outer][n][S]; element m = m1 + n1*m2, output k = k2 + n2*k1:
//   X[k2 + n2 k1] = sum_m1 w_n1^(m1 k1) * [ w_n^(m1 k2) * sum_m2 x[m1 + n1 m2] w_n2^(m2 k2) ].
// PASS 0: lines over (o, m1, c) of length N = n2, stride n1*S: FFT, times
//         w_n^(+-m1 k2), stored in place of the inputs (in -> out).
// PASS 1: lines over (o, k2, c) of length N = n1, stride S: FFT, stored at
//         k2 + n2*k1 (stride n2*S), times `scale`.
// `other` is n1 for PASS 0 and n2 for PASS 1; twn is the table of length n.
template <typename T, int N, bool LF, bool INV, int PASS>
__global__ void __launch_bounds__((P2Launch<T, N, LF, FftV<T, N>::V>::THREADS))
    k_fft_4step(const cx<T> *in, cx<T> *out, size_t S, size_t nlines, size_t other,
                const c128 *__restrict__ tw, const c128 *__restrict__ twn, size_t nbig,
                T scale)
{
    using LP = P2Launch<T, N, LF, FftV<T, N>::V>;
    using P = typename LP::Line;
    constexpr int L = LP::L;
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int tid = threadIdx.x;
    const int l = LF ? tid % L : tid / P::TPL;
    const int t = LF ? tid / L : tid % P::TPL;
    const size_t g = (size_t)blockIdx.x * L + l;
    const bool active = g < nlines;
    const size_t gg = active ? g : 0;
    // input stride: n1*S for pass 0, S for pass 1
    const size_t SI = PASS == 0 ? other * S : S;
    const size_t base = line_base(gg, N, SI);
    cx<T> v[P::E];
#pragma unroll
    for (int u = 0; u < P::E / P::RA_; ++u)
#pragma unroll
        for (int r = 0; r < P::RA_; ++r)
            v[u * P::RA_ + r] =
                active ? in[base + (size_t)P::template in_pos<P::RA_>(t, u, r) * SI] : mk<T>(0, 0);
    // every pass runs the forward radix order; the inverse is the conjugate
    // DFT, taken by conjugating in and out (exact: a sign flip)
    if constexpr (INV) {
#pragma unroll
        for (int e = 0; e < P::E; ++e) v[e].y = -v[e].y;
    }
    P::forward(v, t, lds + l * LP::LSTR, tw, l);
    constexpr int RZ = P::RLAST;
    if (!active) return;
    if constexpr (PASS == 0) {
        const size_t m1 = (gg % (other * S)) / S;
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const int k2 = P::template in_pos<RZ>(t, u, r);
                const size_t m = m1 * (size_t)k2;  // < n1 * n2
                cx<T> a = v[u * RZ + r];
                if constexpr (sizeof(T) == 8) a = a * twn[m];
                else a = cvt<float>(cvt<double>(a) * twn[m]);   // one rounding
                if constexpr (INV) a.y = -a.y;
                out[base + (size_t)k2 * SI] = a;
            }
    } else {
        const size_t row = gg / S, c = gg % S;       // row = o*n2 + k2
        const size_t o = row / other, k2 = row % other;
        const size_t obase = o * nbig * S + k2 * S + c;
#pragma unroll
        for (int u = 0; u < P::E / RZ; ++u)
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const int k1 = P::template in_pos<RZ>(t, u, r);
                cx<T> a = v[u * RZ + r];
                if constexpr (INV) a.y = -a.y;
                out[obase + (size_t)k1 * other * S] = a * scale;
            }
    }
}

// Workgroups are dealt to the 8 XCDs round robin (workgroup b on XCD b % 8).
// For strided line tiles, give each XCD a contiguous range of tiles so the
// 128 B rows neighbouring tiles share stay in one L2.
__device__ __forceinline__ size_t xcd_tile(size_t b, size_t nb)
{
    if (nb % 8) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

// ---------------------------------------------------------------------------
// Generic mixed-radix Stockham for any N whose line fits in LDS twice.  Each
// output of a radix-R pass is its own R-term DFT sum with table twiddles, so
// every factor (2, 3, 5, 7, ..., large primes) takes the same code path.
struct GenFactors {
    int n;
    int nf;
    int f[28];
};

template <typename T>
__global__ void __launch_bounds__(256)
    k_fft_generic(const cx<T> *in, cx<T> *out, size_t S, size_t nlines, int L, GenFactors gf,
                  const c128 *__restrict__ tw, int inverse, T scale)
{
    extern __shared__ unsigned char smem_raw[];
    const int N = gf.n;
    cx<T> *src = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *dst = src + (size_t)L * N;
    const size_t g0 = (size_t)blockIdx.x * L;
    const int total = L * N;
    for (int o = threadIdx.x; o < total; o += blockDim.x) {
        int l, m;
        if (S == 1) {
            l = o / N;
            m = o - l * N;
        } else {
            l = o % L;
            m = o / L;
        }
        size_t g = g0 + l;
        if (g < nlines) src[l * N + m] = in[line_base(g, N, S) + (size_t)m * S];
    }
    __syncthreads();
    int ns = 1;
    for (int p = 0; p < gf.nf; ++p) {
        const int R = gf.f[p];
        const int NR = N / R;
        for (int o = threadIdx.x; o < total; o += blockDim.x) {
            const int l = o / N, pos = o - l * N;
            const int k = pos % ns;
            const int r = (pos / ns) % R;
            const int jb = pos / (ns * R);
            const int j = jb * ns + k;
            const int mstep = k * (N / (ns * R)) + r * NR;
            const cx<T> *s = src + l * N + j;
            cx<T> acc = s[0];
            int e = 0;
            for (int q = 1; q < R; ++q) {
                e += mstep;
                if (e >= N) e -= N;
                const cx<T> w = cvt<T>(tw[e]);
                cx<T> x = s[q * NR];
                acc = acc + (inverse ? mulconj(x, w) : x * w);
            }
            dst[l * N + pos] = acc;
        }
        __syncthreads();
        cx<T> *tmp = src;
        src = dst;
        dst = tmp;
        ns *= R;
    }
    for (int o = threadIdx.x; o < total; o += blockDim.x) {
        int l, m;
        if (S == 1) {
            l = o / N;
            m = o - l * N;
        } else {
            l = o % L;
            m = o / L;
        }
        size_t g = g0 + l;
        if (g < nlines) out[line_base(g, N, S) + (size_t)m * S] = src[l * N + m] * scale;
    }
}

}  // namespace qsim
