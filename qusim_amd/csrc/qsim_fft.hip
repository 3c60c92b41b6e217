// qsim_fft.hip -- FourierTransform{1,2,3}D plans on gfx950 (replaces the
// cuFFT Z2Z plans of src/FourierTransform.cpp:84-128, 336-360, 495-519 and
// the KISS/FFTW CPU transforms).  Power-of-two axes up to 4096 (FP64) / 8192
// (FP32) run the register-resident Stockham kernel; every other length that
// fits twice in LDS runs the generic mixed-radix kernel.
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "qsim_fft_kernels.h"
#include "qsim_fft_plan.h"

namespace qsim {

// ------------------------------------------------------------ twiddle tables
// tw[m] = exp(-2 pi i m / n), m < n, from x87 long double, rounded once to
// double.  Kernels of both precisions read this FP64 table (FP32 kernels
// widen each element, multiply by the FP64 power and round once).
static std::mutex g_tw_mu;
static std::map<std::pair<int, size_t>, void *> g_tw;

int twiddle_table(size_t n, int precision, const void **out)
{
    (void)precision;  // one FP64 table per length serves both precisions
    int dev = 0;
    QSIM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(dev, n);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) {
        *out = it->second;
        return QSIM_OK;
    }
    std::vector<double> h(2 * n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t m = 0; m < n; ++m) {
        long double a = -2.0L * pi * (long double)m / (long double)n;
        h[2 * m] = (double)cosl(a);
        h[2 * m + 1] = (double)sinl(a);
    }
    void *d = nullptr;
    QSIM_HIP(hipMalloc(&d, n * sizeof(c128)));
    QSIM_HIP(hipMemcpy(d, h.data(), n * sizeof(c128), hipMemcpyHostToDevice));
    g_tw[key] = d;
    *out = d;
    return QSIM_OK;
}

bool is_p2_supported(size_t n, int precision)
{
    if (n < 16 || (n & (n - 1))) return false;
    return precision == 64 ? n <= 4096 : n <= 8192;
}

template <typename T, int N, bool LF, bool INV>
static int launch_p2_inst(const void *in, void *out, size_t S, size_t nlines, const void *tw,
                          double scale, hipStream_t s)
{
    using LP = P2Launch<T, N, LF, FftV<T, N>::V>;
    auto fn = k_fft_p2<T, N, LF, INV, false>;
    if constexpr (N >= 128 && N <= 1024) {
        const size_t bytes = nlines * N * sizeof(cx<T>);   // one pass over the array
        if (bytes >= ((size_t)8 << 20) && bytes <= ((size_t)16 << 20))
            fn = k_fft_p2<T, N, LF, INV, true>;
    }
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    size_t blocks = (nlines + LP::L - 1) / LP::L;
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(LP::THREADS), LP::LDS_BYTES, s,
                       (const cx<T> *)in, (cx<T> *)out, S, nlines, (const c128 *)tw, (T)scale);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

template <typename T, int N>
static int launch_p2_n(bool lf, bool inv, const void *in, void *out, size_t S, size_t nlines,
                       const void *tw, double scale, hipStream_t s)
{
    if (lf) {
        return inv ? launch_p2_inst<T, N, true, true>(in, out, S, nlines, tw, scale, s)
                   : launch_p2_inst<T, N, true, false>(in, out, S, nlines, tw, scale, s);
    }
    return inv ? launch_p2_inst<T, N, false, true>(in, out, S, nlines, tw, scale, s)
               : launch_p2_inst<T, N, false, false>(in, out, S, nlines, tw, scale, s);
}

int launch_fft_p2(int precision, size_t n, bool inv, const void *in, void *out, size_t S,
                  size_t nlines, const void *tw, double scale, hipStream_t s)
{
    const bool lf = S > 1;
#define CASE_N(NN)                                                                     \
    case NN:                                                                           \
        return precision == 64                                                         \
                   ? launch_p2_n<double, NN>(lf, inv, in, out, S, nlines, tw, scale, s) \
                   : launch_p2_n<float, NN>(lf, inv, in, out, S, nlines, tw, scale, s);
    switch (n) {
        CASE_N(16)
        CASE_N(32)
        CASE_N(64)
        CASE_N(128)
        CASE_N(256)
        CASE_N(512)
        CASE_N(1024)
        CASE_N(2048)
        CASE_N(4096)
        case 8192:
            if (precision == 32)
                return launch_p2_n<float, 8192>(lf, inv, in, out, S, nlines, tw, scale, s);
            break;
        default:
            break;
    }
#undef CASE_N
    return set_error(QSIM_ERR_UNSUPPORTED, "power-of-two FFT size not instantiated");
}

// Factor order 4, 2, 3, 5, 7, ... (the reference KISS plan, src/kissfft.hh:26-43).
static void factorize(size_t n, GenFactors *gf)
{
    gf->n = (int)n;
    gf->nf = 0;
    size_t m = n, p = 4;
    if (n <= 1) return;
    do {
        while (m % p) {
            if (p == 4) p = 2;
            else if (p == 2) p = 3;
            else p += 2;
            if (p * p > m) p = m;
        }
        m /= p;
        gf->f[gf->nf++] = (int)p;
    } while (m > 1 && gf->nf < 28);
}

// ------------------------------------------------------- four-step passes
// Sub-lengths of the two passes: powers of two 64..4096.
static bool four_sub_ok(size_t m) { return m >= 64 && m <= 4096 && !(m & (m - 1)); }

template <typename T, int N, bool LF, bool INV, int PASS>
static int launch_4s_inst(const void *in, void *out, size_t S, size_t nlines, size_t other,
                          const void *tw, const void *twn, size_t nbig, double scale,
                          hipStream_t s)
{
    using LP = P2Launch<T, N, LF, FftV<T, N>::V>;
    auto fn = k_fft_4step<T, N, LF, INV, PASS>;
    QSIM_TRY(lds_limit((const void *)fn, LP::LDS_BYTES));
    size_t blocks = (nlines + LP::L - 1) / LP::L;
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(LP::THREADS), LP::LDS_BYTES, s,
                       (const cx<T> *)in, (cx<T> *)out, S, nlines, other, (const c128 *)tw,
                       (const c128 *)twn, nbig, (T)scale);
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

// pass 0 reads with stride n1*S (line-fast tiles); pass 1 reads rows (S = 1)
template <typename T, int N>
static int launch_4s_n(int pass, bool inv, const void *in, void *out, size_t S, size_t nlines,
                       size_t other, const void *tw, const void *twn, size_t nbig, double scale,
                       hipStream_t s)
{
    if (pass == 0)
        return inv ? launch_4s_inst<T, N, true, true, 0>(in, out, S, nlines, other, tw, twn,
                                                         nbig, scale, s)
                   : launch_4s_inst<T, N, true, false, 0>(in, out, S, nlines, other, tw, twn,
                                                          nbig, scale, s);
    return inv ? launch_4s_inst<T, N, false, true, 1>(in, out, S, nlines, other, tw, twn, nbig,
                                                      scale, s)
               : launch_4s_inst<T, N, false, false, 1>(in, out, S, nlines, other, tw, twn, nbig,
                                                       scale, s);
}

static int launch_4s(int precision, size_t sub, int pass, bool inv, const void *in, void *out,
                     size_t S, size_t nlines, size_t other, const void *tw, const void *twn,
                     size_t nbig, double scale, hipStream_t s)
{
#define CASE_4S(NN)                                                                          \
    case NN:                                                                                 \
        return precision == 64 ? launch_4s_n<double, NN>(pass, inv, in, out, S, nlines, other, \
                                                         tw, twn, nbig, scale, s)             \
                               : launch_4s_n<float, NN>(pass, inv, in, out, S, nlines, other,  \
                                                        tw, twn, nbig, scale, s);
    switch (sub) {
        CASE_4S(64)
        CASE_4S(128)
        CASE_4S(256)
        CASE_4S(512)
        CASE_4S(1024)
        CASE_4S(2048)
        CASE_4S(4096)
        default:
            break;
    }
#undef CASE_4S
    return set_error(QSIM_ERR_UNSUPPORTED, "four-step sub-length not instantiated");
}

void release_axis(FftAxis *a)
{
    if (a->tmp) (void)hipFree(a->tmp);
    a->tmp = nullptr;
}

int prepare_axis(FftAxis *a, size_t n, size_t S, size_t nlines, int precision)
{
    a->n = n;
    a->S = S;
    a->nlines = nlines;
    a->p2 = is_p2_supported(n, precision);
    QSIM_TRY(twiddle_table(n ? n : 1, precision, &a->tw));
    if (a->p2) return QSIM_OK;
    size_t esz = precision == 64 ? sizeof(c128) : sizeof(c64);
    size_t per_line = 2 * n * esz;
    const size_t lds_max = 160 * 1024;
    if (per_line > lds_max && !(n & (n - 1)) && S == 1) {
        // a long power-of-two innermost axis: n1 = 2^floor(log2(n)/2), n2 = n / n1
        size_t n1 = 1;
        while (n1 * n1 * 4 <= n) n1 *= 2;
        const size_t n2 = n / n1;
        if (four_sub_ok(n1) && four_sub_ok(n2)) {
            a->four = true;
            a->n1 = n1;
            a->n2 = n2;
            QSIM_TRY(twiddle_table(n1, precision, &a->tw1));
            QSIM_TRY(twiddle_table(n2, precision, &a->tw2));
            QSIM_HIP(hipMalloc(&a->tmp, n * nlines * esz));
            return QSIM_OK;
        }
    }
    if (per_line > lds_max)
        return set_error(QSIM_ERR_UNSUPPORTED,
                         "FFT length " + std::to_string(n) +
                             " exceeds the single-pass LDS limit");
    int L = (int)((64 * 1024) / per_line);
    if (L < 1) L = 1;
    if (L > 16) L = 16;
    if (S > 1) {  // a tile of adjacent lines: keep L a divisor of S
        while (L > 1 && (S % L)) --L;
    }
    a->L = L;
    a->lds = (size_t)L * per_line;
    factorize(n, &a->gf);
    return QSIM_OK;
}

int run_axis(const FftAxis &a, int precision, bool inv, const void *in, void *out, double scale,
             hipStream_t s)
{
    // a length-1 axis is a copy/scale: the generic kernel with no factors
    if (a.n == 1 && in == out && scale == 1.0) return QSIM_OK;
    if (a.p2) return launch_fft_p2(precision, a.n, inv, in, out, a.S, a.nlines, a.tw, scale, s);
    if (a.four) {
        // pass 0: nlines*n1 lines of length n2 (stride n1), in -> tmp
        QSIM_TRY(launch_4s(precision, a.n2, 0, inv, in, a.tmp, 1, a.nlines * a.n1, a.n1, a.tw2,
                           a.tw, a.n, 1.0, s));
        // pass 1: nlines*n2 rows of length n1, tmp -> out (transposed store)
        return launch_4s(precision, a.n1, 1, inv, a.tmp, out, 1, a.nlines * a.n2, a.n2, a.tw1,
                         a.tw, a.n, scale, s);
    }
    size_t blocks = (a.nlines + a.L - 1) / a.L;
    if (precision == 64) {
        QSIM_TRY(lds_limit((const void *)k_fft_generic<double>, 160 * 1024));
        hipLaunchKernelGGL(k_fft_generic<double>, dim3((unsigned)blocks), dim3(256), a.lds, s,
                           (const c128 *)in, (c128 *)out, a.S, a.nlines, a.L, a.gf,
                           (const c128 *)a.tw, (int)inv, scale);
    } else {
        QSIM_TRY(lds_limit((const void *)k_fft_generic<float>, 160 * 1024));
        hipLaunchKernelGGL(k_fft_generic<float>, dim3((unsigned)blocks), dim3(256), a.lds, s,
                           (const c64 *)in, (c64 *)out, a.S, a.nlines, a.L, a.gf,
                           (const c128 *)a.tw, (int)inv, (float)scale);
    }
    QSIM_HIP(hipGetLastError());
    return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

struct qsim_fft_plan {
    int rank = 0;
    int inverse = 0;
    int precision = 64;
    size_t dims[3] = {1, 1, 1};
    FftAxis axis[3];
};

extern "C" {

int qsim_hip_fft_plan_create(qsim_fft_plan **plan, int rank, const size_t *dims, int inverse,
                             int precision)
{
    QSIM_CHECK(plan, "null plan out-pointer");
    *plan = nullptr;
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_CHECK(rank >= 1 && rank <= 3, "FFT rank must be 1, 2 or 3");
    QSIM_CHECK(dims, "null dims");
    QSIM_CHECK(precision == 64 || precision == 32, "precision must be 64 or 32");
    size_t total = 1;
    for (int a = 0; a < rank; ++a) {
        QSIM_CHECK(dims[a] >= 1, "FFT dimension must be >= 1");
        total *= dims[a];
    }
    qsim_fft_plan *p = new qsim_fft_plan();
    p->rank = rank;
    p->inverse = inverse ? 1 : 0;
    p->precision = precision;
    for (int a = 0; a < rank; ++a) p->dims[a] = dims[a];
    for (int a = 0; a < rank; ++a) {
        size_t S = 1;
        for (int b = a + 1; b < rank; ++b) S *= dims[b];
        int rc = prepare_axis(&p->axis[a], dims[a], S, total / dims[a], precision);
        if (rc != QSIM_OK) {
            qsim_hip_fft_plan_destroy(p);
            return rc;
        }
    }
    *plan = p;
    return QSIM_OK;
}

int qsim_hip_fft_exec(qsim_fft_plan *p, const void *in, void *out, qsim_stream_t s)
{
    qsim::TraceRange tr_("qsim_hip_fft_exec");
    QSIM_CHECK(p && in && out, "null argument to qsim_hip_fft_exec");
    hipStream_t st = as_stream(s);
    // inner axis first, as the reference transforms (src/FourierTransform.cpp:248-284)
    const void *src = in;
    for (int a = p->rank - 1; a >= 0; --a) {
        if (p->axis[a].n == 1 && src == out) continue;
        QSIM_TRY(run_axis(p->axis[a], p->precision, p->inverse != 0, src, out, 1.0, st));
        src = out;
    }
    if (src != out) {  // every axis had length 1
        size_t bytes = (p->precision == 64 ? sizeof(c128) : sizeof(c64)) * p->dims[0] *
                       p->dims[1] * p->dims[2];
        QSIM_HIP(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, st));
    }
    return QSIM_OK;
}

int qsim_hip_fft_plan_destroy(qsim_fft_plan *p)
{
    QSIM_DEVICE_WAIT_GUARD;
    if (p)
        for (int a = 0; a < 3; ++a) release_axis(&p->axis[a]);
    delete p;
    return QSIM_OK;
}

}  // extern "C"
