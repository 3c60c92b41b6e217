/*
 * qsim_oracle.c -- FP64 CPU restatement of the reference QuSim path
 * (CpuSeq device + KISS FFT).  TEST INFRASTRUCTURE ONLY: the checker for the
 * HIP product path, never part of it.
 *
 * Pinning (each function has an independent check in tests/):
 *   orc_fft          brute-force long-double DFT on small sizes incl. the
 *                    reference's 1234, 123x234, 12x23x34 (applications/QuSimTest/
 *                    TestFFT.cpp:89-100); delta<->constant closed forms;
 *                    round trip ifft(fft(x)) = N x (src/kissfft.hh:70-81);
 *                    golden vectors from the compiled reference KISS FFT.
 *   orc_split*_run   golden trajectories from the compiled reference
 *                    (kissfft.hh + DeviceTemplate.h + SplittingUtils) driven in
 *                    the op order of SplittingMethod1D/2D.cpp; V = 0 closed
 *                    form; norm conservation; tunnelling probability
 *                    0.15937744085 (applications/QuEvol1DPrecision/
 *                    QuSimPrecision.cpp:82); packet velocity 1 (:217-218).
 *     complex V      golden trajectories of the compiled reference's complex
 *                    Exp / MulExp device ops (src/DeviceTemplate.h:1294-1318,
 *                    1468-1496) with an imaginary absorbing layer, 1D and 2D,
 *                    O2 and O4; constant V = -iW gives e^{-WT} times the free
 *                    evolution; the norm never grows when Im V <= 0.
 *     fnes           golden trajectories of Step() with FNES; equals the
 *                    plain run renormalised after every call.
 *   orc_observables  golden values of the compiled reference's Abs2K1D /
 *                    Abs2K2D / Abs2Idx2 / Abs2Mul / Abs2Idx / Norm2 ops;
 *                    plane-wave and Gaussian-packet closed forms of <T>, <V>,
 *                    <x>.
 *   orc_born_run     golden iterates from the compiled reference
 *                    (Perburbation.h BornSeries / PreconditionalBornSeries);
 *                    first-Born closed forms (applications/
 *                    QuScattering{2,3}DPrecision).
 *   orc_xsection*    golden values from the compiled reference Xsection3D;
 *                    first-Born closed forms.
 *   orc_total_xsection*  the first-Born closed form of a weak Gaussian summed
 *                    by the same angular rule, and against the exact integral
 *                    (2D: e^{-x} I0(x); 3D: the reference's weight 4 pi /
 *                    (ntheta npsi) at src/ScatteringSolverImpl.cpp:377 makes
 *                    the total 2 / pi times the integral).
 *   orc_split2d_time, orc_born_time  the run loops above, timed.
 */
#include "qsim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef orc_cplx C;

static inline C cmk(double re, double im) { C z; z.re = re; z.im = im; return z; }
static inline C cadd(C a, C b) { return cmk(a.re + b.re, a.im + b.im); }
static inline C csub(C a, C b) { return cmk(a.re - b.re, a.im - b.im); }
static inline C cmul(C a, C b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline C cscale(C a, double s) { return cmk(a.re * s, a.im * s); }
static inline double cabs2(C a) { return a.im * a.im + a.re * a.re; }
/* exp(z) = e^re (cos im, sin im), as glibc cexp for finite arguments */
static inline C cexp_(C z) { double e = exp(z.re); return cmk(e * cos(z.im), e * sin(z.im)); }
/* a / b, Smith's algorithm (robust complex division) */
static inline C cdiv(C a, C b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return cmk((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    } else {
        double r = b.re / b.im, d = b.re * r + b.im;
        return cmk((a.re * r + a.im) / d, (a.im * r - a.re) / d);
    }
}
/* i * c, reference ExecTrait::IMul (src/DeviceTemplate.h:449-457) */
static inline C cimul(C c) { return cmk(-c.im, c.re); }

static const double kPi = 3.141592653589793238462643383279502884;

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time.  Factor order 4, 2, 3, 5,
 * 7, ... as the reference KISS plan (src/kissfft.hh:26-43); this is our own
 * implementation of the textbook Cooley-Tukey recursion. */

typedef struct {
    size_t n;
    int nstage;
    size_t radix[64];
    size_t rem[64];
    C *tw;       /* tw[j] = exp(sign 2 pi i j / n), j < n */
    int inverse;
} fft_plan;

static fft_plan *plan_create(size_t n, int inverse)
{
    fft_plan *p = (fft_plan *)calloc(1, sizeof(fft_plan));
    p->n = n;
    p->inverse = inverse;
    p->tw = (C *)malloc(sizeof(C) * (n ? n : 1));
    for (size_t j = 0; j < n; ++j) {
        long double a = (inverse ? 2.0L : -2.0L) * 3.141592653589793238462643383279502884L
                        * (long double)j / (long double)n;
        p->tw[j] = cmk((double)cosl(a), (double)sinl(a));
    }
    size_t m = n, f = 4;
    p->nstage = 0;
    if (n <= 1) {
        p->radix[0] = 1; p->rem[0] = 1; p->nstage = 1;
        return p;
    }
    do {
        while (m % f) {
            if (f == 4) f = 2;
            else if (f == 2) f = 3;
            else f += 2;
            if (f * f > m) f = m;
        }
        m /= f;
        p->radix[p->nstage] = f;
        p->rem[p->nstage] = m;
        p->nstage++;
    } while (m > 1);
    return p;
}

static void plan_destroy(fft_plan *p) { if (p) { free(p->tw); free(p); } }

/* Small cache of plans (callers are single-threaded outside the parallel
 * line loops), so a step loop does not rebuild twiddles every transform. */
#define PLAN_CACHE 16
static fft_plan *g_plans[PLAN_CACHE];
static int g_next_plan;
static fft_plan *plan_get(size_t n, int inverse)
{
    for (int i = 0; i < PLAN_CACHE; ++i)
        if (g_plans[i] && g_plans[i]->n == n && g_plans[i]->inverse == inverse) return g_plans[i];
    fft_plan *p = plan_create(n, inverse);
    plan_destroy(g_plans[g_next_plan]);
    g_plans[g_next_plan] = p;
    g_next_plan = (g_next_plan + 1) % PLAN_CACHE;
    return p;
}

/* combine p sub-transforms of length m held consecutively in out[] */
static void butterfly(const fft_plan *pl, C *out, size_t fstride, size_t m, size_t p, C *scratch)
{
    const C *tw = pl->tw;
    const size_t N = pl->n;
    const double s = pl->inverse ? 1.0 : -1.0;
    if (p == 2) {
        for (size_t k = 0; k < m; ++k) {
            C a = out[k], b = cmul(out[k + m], tw[k * fstride]);
            out[k] = cadd(a, b);
            out[k + m] = csub(a, b);
        }
    } else if (p == 4) {
        for (size_t k = 0; k < m; ++k) {
            C t0 = out[k];
            C t1 = cmul(out[k + m], tw[k * fstride]);
            C t2 = cmul(out[k + 2 * m], tw[2 * k * fstride]);
            C t3 = cmul(out[k + 3 * m], tw[3 * k * fstride]);
            C a = cadd(t0, t2), b = csub(t0, t2), c = cadd(t1, t3), d = csub(t1, t3);
            /* s*i*d */
            C sid = cmk(-s * d.im, s * d.re);
            out[k] = cadd(a, c);
            out[k + m] = cadd(b, sid);
            out[k + 2 * m] = csub(a, c);
            out[k + 3 * m] = csub(b, sid);
        }
    } else if (p == 3) {
        const double h = 0.86602540378443864676372317075294; /* sqrt(3)/2 */
        for (size_t k = 0; k < m; ++k) {
            C t0 = out[k];
            C t1 = cmul(out[k + m], tw[k * fstride]);
            C t2 = cmul(out[k + 2 * m], tw[2 * k * fstride]);
            C u = cadd(t1, t2), v = csub(t1, t2);
            C base = cmk(t0.re - 0.5 * u.re, t0.im - 0.5 * u.im);
            C rot = cmk(-s * h * v.im, s * h * v.re); /* s*i*h*v */
            out[k] = cadd(t0, u);
            out[k + m] = cadd(base, rot);
            out[k + 2 * m] = csub(base, rot);
        }
    } else {
        for (size_t k = 0; k < m; ++k) {
            for (size_t q = 0; q < p; ++q)
                scratch[q] = cmul(out[k + q * m], tw[(q * k * fstride) % N]);
            for (size_t r = 0; r < p; ++r) {
                C acc = scratch[0];
                for (size_t q = 1; q < p; ++q)
                    acc = cadd(acc, cmul(scratch[q], tw[(q * r * m * fstride) % N]));
                out[k + r * m] = acc;
            }
        }
    }
}

static void fft_rec(const fft_plan *pl, const C *in, size_t istride, C *out,
                    int stage, size_t fstride, C *scratch)
{
    size_t p = pl->radix[stage], m = pl->rem[stage];
    if (m == 1) {
        for (size_t q = 0; q < p; ++q) out[q] = in[q * fstride * istride];
    } else {
        for (size_t q = 0; q < p; ++q)
            fft_rec(pl, in + q * fstride * istride, istride, out + q * m, stage + 1,
                    fstride * p, scratch);
    }
    if (p > 1) butterfly(pl, out, fstride, m, p, scratch);
}

/* one transform of a strided line: in[k*stride] -> out[k*stride] */
static void fft_line(const fft_plan *pl, const C *in, C *out, size_t stride, C *tmpin, C *tmpout,
                     C *scratch)
{
    size_t n = pl->n;
    for (size_t k = 0; k < n; ++k) tmpin[k] = in[k * stride];
    fft_rec(pl, tmpin, 1, tmpout, 0, 1, scratch);
    for (size_t k = 0; k < n; ++k) out[k * stride] = tmpout[k];
}

static size_t max_radix(const fft_plan *p)
{
    size_t r = 1;
    for (int i = 0; i < p->nstage; ++i) if (p->radix[i] > r) r = p->radix[i];
    return r;
}

/* Transform every line of length dims[axis] of the row-major array, in -> out. */
static void fft_axis(int rank, const size_t *dims, int axis, int inverse, const C *in, C *out,
                     int threads)
{
    size_t n = dims[axis], stride = 1, outer = 1;
    for (int d = axis + 1; d < rank; ++d) stride *= dims[d];
    for (int d = 0; d < axis; ++d) outer *= dims[d];
    size_t nlines = outer * stride;
    fft_plan *pl = plan_get(n, inverse);
    size_t mr = max_radix(pl);
    (void)threads;
#pragma omp parallel num_threads(threads > 0 ? threads : 1) if (threads > 1)
    {
        C *ti = (C *)malloc(sizeof(C) * n), *to = (C *)malloc(sizeof(C) * n);
        C *sc = (C *)malloc(sizeof(C) * mr);
#pragma omp for schedule(static)
        for (long long l = 0; l < (long long)nlines; ++l) {
            size_t o = (size_t)l / stride, c = (size_t)l % stride;
            size_t base = o * n * stride + c;
            fft_line(pl, in + base, out + base, stride, ti, to, sc);
        }
        free(ti); free(to); free(sc);
    }
}

int orc_fft(int rank, const size_t *dims, int inverse, const orc_cplx *in, orc_cplx *out,
            int threads)
{
    if (rank < 1 || rank > 3) return -1;
    /* inner axis first, as the reference KISS 2D/3D transforms */
    fft_axis(rank, dims, rank - 1, inverse, in, out, threads);
    for (int a = rank - 2; a >= 0; --a) fft_axis(rank, dims, a, inverse, out, out, threads);
    return 0;
}

int orc_dft_bruteforce(size_t n, int inverse, const orc_cplx *in, orc_cplx *out)
{
    long double s = inverse ? 1.0L : -1.0L;
    for (size_t k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (size_t j = 0; j < n; ++j) {
            size_t m = (j * k) % n;
            long double a = s * 2.0L * 3.141592653589793238462643383279502884L * (long double)m
                            / (long double)n;
            long double c = cosl(a), sn = sinl(a);
            re += in[j].re * c - in[j].im * sn;
            im += in[j].re * sn + in[j].im * c;
        }
        out[k] = cmk((double)re, (double)im);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Device-op restatements (src/DeviceTemplate.h) */

double orc_norm2(const orc_cplx *a, size_t n)
{
    double s = 0; /* plain left-to-right sum, as std::transform_reduce seq */
    for (size_t i = 0; i < n; ++i) s += cabs2(a[i]);
    return s;
}

/* fold_half, src/DeviceTemplate.h:544-547 */
static inline size_t fold_half(size_t i, size_t n) { return (i > n / 2) ? (n - i) : i; }

/* psi *= exp(alpha * V), op_MulExpMul (src/DeviceTemplate.h:617-628) */
static void mul_exp_v(C *psi, C alpha, const C *V, int complex_v, size_t n, int threads)
{
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) if (threads > 1) schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        C z = complex_v ? cmul(alpha, V[i]) : cscale(alpha, V[i].re);
        psi[i] = cmul(psi[i], cexp_(z));
    }
}

/* to = exp(alpha * V), op_ExpMul (src/DeviceTemplate.h:604-615) */
static void exp_v(C *to, C alpha, const C *V, int complex_v, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        C z = complex_v ? cmul(alpha, V[i]) : cscale(alpha, V[i].re);
        to[i] = cexp_(z);
    }
}

static void mul_inplace(C *to, const C *from, size_t n, int threads)
{
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) if (threads > 1) schedule(static)
    for (long long i = 0; i < (long long)n; ++i) to[i] = cmul(from[i], to[i]);
}

static void scale_inplace(C *a, double s, size_t n, int threads)
{
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) if (threads > 1) schedule(static)
    for (long long i = 0; i < (long long)n; ++i) a[i] = cscale(a[i], s);
}

/* ------------------------------------------------------------------------ */
/* Split-operator constants, src/SplittingUtils.cpp:3-13 */
static double sC1, sC2, sD1, sD2;
static void split_constants(void)
{
    double t = pow(2, 1 / 3.0);
    sC1 = 1 / (2 * (2 - t));
    sC2 = (1 - t) / (2 * (2 - t));
    sD1 = 1 / (2 - t);
    sD2 = -t / (2 - t);
}

/* f = -I / hbar * dt * t, evaluated in the reference's order
 * (src/SplittingMethod1D.cpp:86, src/SplittingMethod2D.cpp:43) */
static C phase_v(double hbar, C dt, double t)
{
    C mi = cmk(-0.0, -1.0);
    C a = cmk(mi.re / hbar, mi.im / hbar);
    C b = cmul(a, dt);
    return cscale(b, t);
}

/* ---- 1D ---- */
typedef struct {
    size_t n;
    double dx, mass, hbar;
    C dt;
    int boundary;
    const C *V;
    int complex_v;
    C *fft_buf, *iw_psi, *iw_kpsi;
    int threads;
} split1d;

/* SplittingMethod1D::ExpV, src/SplittingMethod1D.cpp:84-89 */
static void s1_expv(split1d *s, C *psi, double t)
{
    mul_exp_v(psi, phase_v(s->hbar, s->dt, t), s->V, s->complex_v, s->n, s->threads);
}

/* SplittingMethod1D::ExpT, src/SplittingMethod1D.cpp:40-81 */
static void s1_expt(split1d *s, C *psi, double tt)
{
    size_t N = s->n;
    if (s->boundary == 0) {
        orc_fft(1, &N, 0, psi, s->fft_buf, 1);
        double Dk = 2 * kPi / (s->dx * N);
        double DT = (Dk * s->hbar) * (Dk * s->hbar) / (2 * s->mass);
        /* f = -I * DT * fDt * tt / fHbar */
        C f = cmk(-0.0 * DT, -1.0 * DT);
        f = cmul(f, s->dt);
        f = cscale(f, tt);
        f = cmk(f.re / s->hbar, f.im / s->hbar);
        for (size_t i = 0; i < N; ++i) {
            double k = (double)fold_half(i, N);
            s->fft_buf[i] = cmul(s->fft_buf[i], cexp_(cscale(f, k * k)));
        }
        orc_fft(1, &N, 1, s->fft_buf, psi, 1);
        scale_inplace(psi, 1. / N, N, s->threads);
    } else {
        size_t N2 = 2 * N;
        C *a = s->iw_psi, *b = s->iw_kpsi;
        memcpy(a, psi, sizeof(C) * N);
        a[0] = cmk(0, 0);
        a[N] = cmk(0, 0);
        for (size_t m = 0; m + 1 < N; ++m) a[N2 - 1 - m] = cmk(-a[1 + m].re, -a[1 + m].im);
        orc_fft(1, &N2, 1, a, b, 1);
        double Dp = s->hbar * kPi / (s->dx * N);
        double DT = Dp * Dp / (2 * s->mass);
        C f = cmk(-0.0 * DT, -1.0 * DT);
        f = cmul(f, s->dt);
        f = cscale(f, tt);
        f = cmk(f.re / s->hbar, f.im / s->hbar);
        for (size_t i = 0; i < N; ++i) {
            double k = (double)i;
            b[i] = cmul(b[i], cexp_(cscale(f, k * k)));
        }
        b[0] = cmk(0, 0);
        b[N] = cmk(0, 0);
        for (size_t m = 0; m + 1 < N; ++m) b[N2 - 1 - m] = cmk(-b[1 + m].re, -b[1 + m].im);
        orc_fft(1, &N2, 1, b, a, 1);
        double sc = -1. / (2. * N);
        for (size_t i = 0; i < N; ++i) psi[i] = cscale(a[i], sc);
    }
}

int orc_split1d_run(size_t n, double x0, double x1, double dt_re, double dt_im, double mass,
                    double hbar, int method, int boundary, const orc_cplx *V, int complex_v,
                    orc_cplx *psi, size_t nsteps, size_t batch, int fnes, int threads)
{
    if (method != 2 && method != 4) return -1;
    split_constants();
    split1d s;
    memset(&s, 0, sizeof s);
    s.n = n;
    s.dx = (x1 - x0) / n;
    s.mass = mass;
    s.hbar = hbar;
    s.dt = cmk(dt_re, dt_im);
    s.boundary = boundary;
    s.V = V;
    s.complex_v = complex_v;
    s.threads = threads;
    s.fft_buf = (C *)malloc(sizeof(C) * n);
    s.iw_psi = (C *)malloc(sizeof(C) * 2 * n);
    s.iw_kpsi = (C *)malloc(sizeof(C) * 2 * n);
    for (size_t st = 0; st < nsteps; ++st) {
        /* QuUpdatePsi, src/SplittingUtils.h:14-64 (fold_head_tail = false) */
        for (size_t b = 0; b < batch; ++b) {
            if (method == 2) {
                s1_expv(&s, psi, 0.5);
                s1_expt(&s, psi, 1.0);
                s1_expv(&s, psi, 0.5);
            } else {
                s1_expv(&s, psi, sC1);
                s1_expt(&s, psi, sD1);
                s1_expv(&s, psi, sC2);
                s1_expt(&s, psi, sD2);
                s1_expv(&s, psi, sC2);
                s1_expt(&s, psi, sD1);
                s1_expv(&s, psi, sC1);
            }
        }
        if (fnes) { /* QuEvolverImpl::Step, src/EvolverImpl.cpp:60-63 */
            double sc = 1.0 / sqrt(orc_norm2(psi, n) * s.dx);
            scale_inplace(psi, sc, n, threads);
        }
    }
    free(s.fft_buf); free(s.iw_psi); free(s.iw_kpsi);
    return 0;
}

double orc_norm2_range_1d(const orc_cplx *psi, size_t n, double x0, double x1, double a,
                          double b)
{
    /* QuEvolver1DImpl::Norm2(x0, x1), src/EvolverImpl1D.cpp:89-101 */
    double dx = (x1 - x0) / n;
    long long i1 = (long long)((a - x0) / dx), i2 = (long long)((b - x0) / dx);
    if (i1 <= 0) i1 = 0;
    if (i2 >= (long long)n) i2 = (long long)n;
    if (i1 >= i2) return 0;
    return orc_norm2(psi + i1, (size_t)(i2 - i1)) * dx;
}

/* Evolver observables, out[0..3] = KinEn, PotEn, Norm2 (times dv), Xavg (1D).
 *   KinEn 1D periodic: SplittingMethod1D::CalKinEn (src/SplittingMethod1D.cpp:94-110):
 *     sum |FFT psi|^2 fold(i)^2 DT / sum |FFT psi|^2, DT = Dk^2 / 2m (no hbar, as there)
 *   KinEn 1D wall (:112-128): odd extension to 2N, inverse 2N DFT,
 *     sum_{i<N} |b_i|^2 i^2 DT / sum_{i<N} |b_i|^2, DT = (hbar pi/(dx N))^2 / 2m
 *   KinEn 2D: SplittingMethod2D::CalKinEn (src/SplittingMethod2D.cpp:124-141)
 *   PotEn:    QuEvolverImpl::CalPotEn (src/EvolverImpl.cpp:45-49): sum |psi|^2 V dv / Norm2
 *   Xavg:     QuEvolver1DImpl::Xavg (src/EvolverImpl1D.cpp:82-87): x0 + dx sum|psi|^2 i / sum|psi|^2
 */
int orc_observables(int dim, size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                    double mass, double hbar, int boundary, const orc_cplx *psi,
                    const double *V, double *out)
{
    if (dim != 1 && dim != 2) return -1;
    const size_t n = dim == 1 ? nx : nx * ny;
    const double dx = (x1 - x0) / nx, dy = dim == 1 ? 1.0 : (y1 - y0) / ny;
    double kin = 0;
    if (dim == 1 && boundary == 0) {
        C *ft = (C *)malloc(sizeof(C) * n);
        orc_fft(1, &n, 0, psi, ft, 1);
        double Dk = 2 * kPi / (dx * n), DT = Dk * Dk / (2 * mass);
        double s = 0;
        for (size_t i = 0; i < n; ++i) {
            double k = (double)fold_half(i, n);
            s += cabs2(ft[i]) * (k * k) * DT;
        }
        kin = s / orc_norm2(ft, n);
        free(ft);
    } else if (dim == 1) {
        size_t N2 = 2 * n;
        C *a = (C *)malloc(sizeof(C) * N2), *b = (C *)malloc(sizeof(C) * N2);
        memcpy(a, psi, sizeof(C) * n);
        a[0] = cmk(0, 0);
        a[n] = cmk(0, 0);
        for (size_t m = 0; m + 1 < n; ++m) a[N2 - 1 - m] = cmk(-a[1 + m].re, -a[1 + m].im);
        orc_fft(1, &N2, 1, a, b, 1);
        double Dk = kPi / (dx * n), DT = (Dk * hbar) * (Dk * hbar) / (2 * mass);
        double s = 0;
        for (size_t i = 0; i < n; ++i) s += cabs2(b[i]) * ((double)i * (double)i);
        kin = s * DT / orc_norm2(b, n);
        free(a);
        free(b);
    } else {
        C *ft = (C *)malloc(sizeof(C) * n);
        size_t dims[2] = {nx, ny};
        orc_fft(2, dims, 0, psi, ft, 1);
        double Dkx = 2 * kPi / (dx * nx), Dky = 2 * kPi / (dy * ny);
        double DTx = (Dkx * hbar) * (Dkx * hbar) / (2 * mass);
        double DTy = (Dky * hbar) * (Dky * hbar) / (2 * mass);
        double s = 0;
        for (size_t i = 0; i < nx; ++i)
            for (size_t j = 0; j < ny; ++j) {
                double ki = (double)fold_half(i, nx), kj = (double)fold_half(j, ny);
                s += cabs2(ft[i * ny + j]) * (ki * ki * DTx + kj * kj * DTy);
            }
        kin = s / orc_norm2(ft, n);
        free(ft);
    }
    const double dv = dx * dy;
    const double nrm = orc_norm2(psi, n);
    double pot = 0, idx = 0;
    for (size_t i = 0; i < n; ++i) {
        pot += cabs2(psi[i]) * V[i];
        idx += cabs2(psi[i]) * (double)i;
    }
    out[0] = kin;
    out[1] = pot * dv / (nrm * dv);
    out[2] = nrm * dv;
    out[3] = dim == 1 ? x0 + dx * (idx / nrm) : 0;
    return 0;
}

/* ---- 2D ---- */
typedef struct {
    size_t nx, ny, n;
    C *ev_half, *ev_c1, *ev_c2, *et_1, *et_d1, *et_d2, *ft;
    int threads;      /* pointwise ops */
    int fft_threads;  /* lines of the 2D FFT */
} split2d;

/* SplittingMethod2D::InitExpT, src/SplittingMethod2D.cpp:56-79 */
static void s2_make_et(C *t, size_t nx, size_t ny, double DTx, double DTy, C alpha)
{
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < ny; ++j) {
            double fi = (double)fold_half(i, nx), fj = (double)fold_half(j, ny);
            double e = fi * fi * DTx + fj * fj * DTy;
            t[i * ny + j] = cexp_(cscale(alpha, e));
        }
}

static void s2_fft(split2d *s, const C *in, C *out, int inverse)
{
    size_t d[2] = {s->nx, s->ny};
    orc_fft(2, d, inverse, in, out, s->fft_threads);
}

/* SplittingMethod2D::ExpT, src/SplittingMethod2D.cpp:96-121 */
static void s2_expt(split2d *s, C *psi, const C *et)
{
    s2_fft(s, psi, s->ft, 0);
    mul_inplace(s->ft, et, s->n, s->threads);
    s2_fft(s, s->ft, psi, 1);
    scale_inplace(psi, 1. / s->n, s->n, s->threads);
}

static void s2_init(split2d *s, size_t nx, size_t ny, double x0, double x1, double y0,
                    double y1, C dt, double mass, double hbar, int method, const C *V,
                    int complex_v, int threads)
{
    split_constants();
    memset(s, 0, sizeof *s);
    s->nx = nx; s->ny = ny; s->n = nx * ny; s->threads = threads; s->fft_threads = threads;
    double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
    size_t n = s->n;
    s->ft = (C *)malloc(sizeof(C) * n);
    /* InitExpV, src/SplittingMethod2D.cpp:38-54 */
    if (method == 2) {
        s->ev_half = (C *)malloc(sizeof(C) * n);
        exp_v(s->ev_half, phase_v(hbar, dt, 0.5), V, complex_v, n);
    } else {
        s->ev_c1 = (C *)malloc(sizeof(C) * n);
        s->ev_c2 = (C *)malloc(sizeof(C) * n);
        exp_v(s->ev_c1, phase_v(hbar, dt, sC1), V, complex_v, n);
        exp_v(s->ev_c2, phase_v(hbar, dt, sC2), V, complex_v, n);
    }
    double Dkx = 2 * kPi / (dx * nx), Dky = 2 * kPi / (dy * ny);
    double DTx = (Dkx * hbar) * (Dkx * hbar) / (2 * mass);
    double DTy = (Dky * hbar) * (Dky * hbar) / (2 * mass);
    /* alpha = -I * fDt / fHbar */
    C alpha = cmul(cmk(-0.0, -1.0), dt);
    alpha = cmk(alpha.re / hbar, alpha.im / hbar);
    if (method == 2) {
        s->et_1 = (C *)malloc(sizeof(C) * n);
        s2_make_et(s->et_1, nx, ny, DTx, DTy, alpha);
    } else {
        s->et_d1 = (C *)malloc(sizeof(C) * n);
        s->et_d2 = (C *)malloc(sizeof(C) * n);
        s2_make_et(s->et_d1, nx, ny, DTx, DTy, cscale(alpha, sD1));
        s2_make_et(s->et_d2, nx, ny, DTx, DTy, cscale(alpha, sD2));
    }
}

static void s2_free(split2d *s)
{
    free(s->ev_half); free(s->ev_c1); free(s->ev_c2);
    free(s->et_1); free(s->et_d1); free(s->et_d2); free(s->ft);
}

static void s2_step(split2d *s, C *psi, int method)
{
    /* QuUpdatePsi body, src/SplittingUtils.h:28-58 */
    if (method == 2) {
        mul_inplace(psi, s->ev_half, s->n, s->threads);
        s2_expt(s, psi, s->et_1);
        mul_inplace(psi, s->ev_half, s->n, s->threads);
    } else {
        mul_inplace(psi, s->ev_c1, s->n, s->threads);
        s2_expt(s, psi, s->et_d1);
        mul_inplace(psi, s->ev_c2, s->n, s->threads);
        s2_expt(s, psi, s->et_d2);
        mul_inplace(psi, s->ev_c2, s->n, s->threads);
        s2_expt(s, psi, s->et_d1);
        mul_inplace(psi, s->ev_c1, s->n, s->threads);
    }
}

int orc_split2d_run(size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                    double dt_re, double dt_im, double mass, double hbar, int method,
                    const orc_cplx *V, int complex_v, orc_cplx *psi, size_t nsteps,
                    size_t batch, int fnes, int threads)
{
    if (method != 2 && method != 4) return -1;
    split2d s;
    s2_init(&s, nx, ny, x0, x1, y0, y1, cmk(dt_re, dt_im), mass, hbar, method, V, complex_v,
            threads);
    double dv = ((x1 - x0) / nx) * ((y1 - y0) / ny);
    for (size_t st = 0; st < nsteps; ++st) {
        for (size_t b = 0; b < batch; ++b) s2_step(&s, psi, method);
        if (fnes) {
            double sc = 1.0 / sqrt(orc_norm2(psi, s.n) * dv);
            scale_inplace(psi, sc, s.n, threads);
        }
    }
    s2_free(&s);
    return 0;
}

static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0;
#endif
}

double orc_split2d_time(size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                        double dt, double mass, double hbar, const orc_cplx *V, int complex_v,
                        orc_cplx *psi, size_t nsteps, int threads, int fft_threads)
{
    split2d s;
    s2_init(&s, nx, ny, x0, x1, y0, y1, cmk(dt, 0), mass, hbar, 2, V, complex_v, threads);
    s.fft_threads = fft_threads;
    double t0 = now_s();
    for (size_t st = 0; st < nsteps; ++st) s2_step(&s, psi, 2);
    double t1 = now_s();
    s2_free(&s);
    return t1 - t0;
}

/* ------------------------------------------------------------------------ */
/* Born series */

/* cal_v15, src/Perburbation.h:45-51 */
static double cal_v15(double x, double k0, double e0)
{
    x *= k0;
    return (-0.8 * e0) * ((1 - x * x / (15. * 15)) > 0
                              ? exp(-0.015 * x * x * 1. / sqrt(1 - x * x / (15. * 15)))
                              : 0);
}

/* fold_half of src/Perburbation.h:30-33 */
static inline size_t pfold(size_t i, size_t n) { return (i < n / 2) ? i : n - i; }

/* CalDp2Div2M, src/Perburbation.h:16-22 */
static double dp2div2m(double dx, size_t n, double hbar, double mass)
{
    double Dpx = 2 * kPi / (dx * n) * hbar;
    return Dpx * Dpx / (2 * mass);
}

int orc_plane_wave(int dim, const size_t *n, const double *lo, const double *hi, double energy,
                   const double *direction, double mass, double hbar, orc_cplx *psi0)
{
    double k0 = sqrt(2 * energy * mass) / hbar;
    double d[3] = {0, 0, 0}, nm2 = 0;
    for (int a = 0; a < dim; ++a) nm2 += direction[a] * direction[a];
    double nm = 1 / sqrt(nm2);
    for (int a = 0; a < dim; ++a) d[a] = direction[a] * nm * k0;
    double dx[3];
    for (int a = 0; a < dim; ++a) dx[a] = (hi[a] - lo[a]) / n[a];
    if (dim == 2) {
        for (size_t i = 0; i < n[0]; ++i)
            for (size_t j = 0; j < n[1]; ++j) {
                double x = lo[0] + dx[0] * i, y = lo[1] + dx[1] * j;
                double ph = d[0] * x + d[1] * y;
                psi0[i * n[1] + j] = cmk(cos(ph), sin(ph));
            }
    } else {
        for (size_t i = 0; i < n[0]; ++i)
            for (size_t j = 0; j < n[1]; ++j)
                for (size_t k = 0; k < n[2]; ++k) {
                    double x = lo[0] + dx[0] * i, y = lo[1] + dx[1] * j,
                           z = lo[2] + dx[2] * k;
                    double ph = d[0] * x + d[1] * y + d[2] * z;
                    psi0[(i * n[1] + j) * n[2] + k] = cmk(cos(ph), sin(ph));
                }
    }
    return 0;
}

/* G0 multiply, op_G02D / op_G03D (src/DeviceTemplate.h:905-976) */
static void g0_mul(int dim, const size_t *n, C *psik, double E, double eps, const double *a,
                   int threads)
{
    size_t tot = n[0] * n[1] * (dim == 3 ? n[2] : 1);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) if (threads > 1) schedule(static)
    for (long long idx = 0; idx < (long long)tot; ++idx) {
        double t;
        if (dim == 2) {
            size_t i = (size_t)idx / n[1], j = (size_t)idx % n[1];
            double fi = (double)fold_half(i, n[0]), fj = (double)fold_half(j, n[1]);
            t = fi * fi * a[0] + fj * fj * a[1];
        } else {
            size_t k = (size_t)idx % n[2], i2 = (size_t)idx / n[2];
            size_t j = i2 % n[1], i = i2 / n[1];
            double fi = (double)fold_half(i, n[0]), fj = (double)fold_half(j, n[1]),
                   fk = (double)fold_half(k, n[2]);
            t = fi * fi * a[0] + fj * fj * a[1] + fk * fk * a[2];
        }
        psik[idx] = cdiv(psik[idx], cmk(E - t, eps));
    }
}

static int born_impl(int dim, const size_t *n, const double *lo, const double *hi,
                     double energy, const double *direction, double mass, double hbar,
                     int preconditional, int preconditioner, double slow, int order,
                     const orc_cplx *Vin, orc_cplx *dpsi, orc_cplx *v_eff, double *eps_inout,
                     double *delta_norm, int ncalls, int threads, int fft_threads,
                     double *loop_seconds);

int orc_born_run(int dim, const size_t *n, const double *lo, const double *hi, double energy,
                 const double *direction, double mass, double hbar, int preconditional,
                 int preconditioner, double slow, int order, const orc_cplx *Vin,
                 orc_cplx *dpsi, orc_cplx *v_eff, double *eps_inout, double *delta_norm,
                 int ncalls, int threads)
{
    /* the pointwise loops stay serial here (bitwise the results of the
     * reference's sequential ops); the FFT lines use `threads` */
    return born_impl(dim, n, lo, hi, energy, direction, mass, hbar, preconditional,
                     preconditioner, slow, order, Vin, dpsi, v_eff, eps_inout, delta_norm,
                     ncalls, 1, threads, NULL);
}

double orc_born_time(int dim, const size_t *n, const double *lo, const double *hi,
                     double energy, const double *direction, double mass, double hbar,
                     int preconditional, int preconditioner, double slow, int iters,
                     const orc_cplx *Vin, int threads, int fft_threads)
{
    size_t tot = n[0] * n[1] * (dim == 3 ? n[2] : 1);
    C *dpsi = (C *)calloc(tot, sizeof(C));
    double eps = 0, secs = -1;
    if (born_impl(dim, n, lo, hi, energy, direction, mass, hbar, preconditional,
                  preconditioner, slow, iters, Vin, dpsi, NULL, &eps, NULL, 1, threads,
                  fft_threads, &secs) != 0)
        secs = -1;
    free(dpsi);
    return secs;
}

static int born_impl(int dim, const size_t *n, const double *lo, const double *hi,
                     double energy, const double *direction, double mass, double hbar,
                     int preconditional, int preconditioner, double slow, int order,
                     const orc_cplx *Vin, orc_cplx *dpsi, orc_cplx *v_eff, double *eps_inout,
                     double *delta_norm, int ncalls, int threads, int fft_threads,
                     double *loop_seconds)
{
    const int nt = threads > 0 ? threads : 1;
    if (dim != 2 && dim != 3) return -1;
    size_t tot = n[0] * n[1] * (dim == 3 ? n[2] : 1);
    double dx[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) dx[a] = (hi[a] - lo[a]) / n[a];
    C *V = (C *)malloc(sizeof(C) * tot), *psi0 = (C *)malloc(sizeof(C) * tot);
    C *psik = (C *)malloc(sizeof(C) * tot), *tmp = (C *)malloc(sizeof(C) * tot);
    C *last = (C *)malloc(sizeof(C) * tot);
    memcpy(V, Vin, sizeof(C) * tot);
    orc_plane_wave(dim, n, lo, hi, energy, direction, mass, hbar, psi0);
    double eps = eps_inout ? *eps_inout : 0;
    double k0 = sqrt(2 * mass * energy) / hbar;
    if (preconditional) {
        /* GaussMAsbLayer2D/3D, src/Perburbation.h:131-177 */
        if (dim == 2) {
            for (size_t i = 0; i < n[0]; ++i)
                for (size_t j = 0; j < n[1]; ++j) {
                    double x = pfold(i, n[0]) * dx[0], y = pfold(j, n[1]) * dx[1];
                    size_t idx = i * n[1] + j;
                    V[idx].im = V[idx].im + (cal_v15(x, k0, energy) + cal_v15(y, k0, energy));
                }
        } else {
            for (size_t i = 0; i < n[0]; ++i)
                for (size_t j = 0; j < n[1]; ++j)
                    for (size_t k = 0; k < n[2]; ++k) {
                        double x = pfold(i, n[0]) * dx[0], y = pfold(j, n[1]) * dx[1],
                               z = pfold(k, n[2]) * dx[2];
                        size_t idx = (i * n[1] + j) * n[2] + k;
                        V[idx].im = V[idx].im + (cal_v15(x, k0, energy) + cal_v15(y, k0, energy)
                                                 + cal_v15(z, k0, energy));
                    }
        }
        /* GetEpsilon, src/Perburbation.h:329-344 */
        double minEps;
        if (preconditioner == 0) {
            double A = 0, A2 = 0;
            for (size_t i = 0; i < tot; ++i) A += V[i].im;
            A2 = orc_norm2(V, tot);
            minEps = 0.5 * A2 / (-A);
        } else {
            double m = 0;
            for (size_t i = 0; i < tot; ++i) {
                double v = cabs2(V[i]);
                m = m >= v ? m : v;
            }
            minEps = sqrt(m);
        }
        if (eps < minEps) eps = minEps;
    }
    double a[3];
    for (int d = 0; d < dim; ++d) a[d] = dp2div2m(dx[d], n[d], hbar, mass);
    size_t dims[3] = {n[0], n[1], dim == 3 ? n[2] : 1};
    double rn = 1. / (double)tot;
    double dnorm = 0;
    const double t_start = now_s();
    for (int call = 0; call < ncalls; ++call) {
        memcpy(last, dpsi, sizeof(C) * tot);
        for (int it = 0; it < order; ++it) {
            if (preconditional) {
                /* PreconditionalBornSeries::Update, src/Perburbation.h:350-423 */
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
                for (long long i = 0; i < (long long)tot; ++i) /* op_Add3 */
                    tmp[i] = cadd(cmul(cmk(V[i].re, V[i].im + eps), dpsi[i]),
                                  cscale(psi0[i], V[i].re));
                orc_fft(dim, dims, 0, tmp, psik, fft_threads);
                g0_mul(dim, n, psik, energy, eps, a, nt);
                orc_fft(dim, dims, 1, psik, tmp, fft_threads);
                if (preconditioner == 0 && slow == 1.) {
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
                    for (long long i = 0; i < (long long)tot; ++i) dpsi[i] = cscale(tmp[i], rn);
                } else {
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
                    for (long long i = 0; i < (long long)tot; ++i) tmp[i] = cscale(tmp[i], rn);
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
                    for (long long i = 0; i < (long long)tot; ++i) {
                        C g;
                        if (preconditioner == 0) {
                            dpsi[i] = cadd(cscale(dpsi[i], 1 - slow), cscale(tmp[i], slow));
                            continue;
                        } else if (preconditioner == 1) {      /* op_Vellekoop */
                            C iv = cimul(V[i]);
                            g = cmk(1. - iv.re / eps, -iv.im / eps);
                        } else if (preconditioner == 2) {      /* op_Hao1 */
                            C iv = cimul(cmk(V[i].re, -V[i].im));
                            g = cmk(1. - iv.re / eps, -iv.im / eps);
                        } else {                               /* op_Hao2 */
                            C iv = cimul(V[i]);
                            g = cdiv(cmk(2., 0), cmk(1. + iv.re / eps, iv.im / eps));
                        }
                        g = cscale(g, slow);
                        dpsi[i] = cadd(cmul(cmk(1. - g.re, -g.im), dpsi[i]), cmul(g, tmp[i]));
                    }
                }
            } else {
                /* BornSeries::Update2D/3D, src/Perburbation.h:257-309 */
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
                for (long long i = 0; i < (long long)tot; ++i)
                    dpsi[i] = cmul(cadd(dpsi[i], psi0[i]), V[i]);
                orc_fft(dim, dims, 0, dpsi, psik, fft_threads);
                g0_mul(dim, n, psik, energy, eps, a, nt);
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
                for (long long i = 0; i < (long long)tot; ++i) psik[i] = cscale(psik[i], rn);
                orc_fft(dim, dims, 1, psik, dpsi, fft_threads);
            }
        }
        if (loop_seconds) *loop_seconds = now_s() - t_start;
        double num = 0, den = 0;
        for (size_t i = 0; i < tot; ++i) num += cabs2(csub(dpsi[i], last[i]));
        den = orc_norm2(dpsi, tot);
        dnorm = sqrt(num / den);
    }
    if (v_eff) memcpy(v_eff, V, sizeof(C) * tot);
    if (eps_inout) *eps_inout = eps;
    if (delta_norm) *delta_norm = dnorm;
    free(V); free(psi0); free(psik); free(tmp); free(last);
    return 0;
}

double orc_xsection2d(const size_t *n, const double *lo, const double *hi, double energy,
                      double mass, double hbar, const orc_cplx *psi0, const orc_cplx *dpsi,
                      const orc_cplx *V, double cosx, double cosy)
{
    /* ScatteringSolver2DImpl::ComputeXSection, src/ScatteringSolverImpl.cpp:168-196 */
    double nm = 1 / sqrt(cosx * cosx + cosy * cosy);
    cosx *= nm;
    cosy *= nm;
    double k0 = sqrt(2 * energy * mass) / hbar;
    double kx = k0 * cosx, ky = k0 * cosy;
    double dx = (hi[0] - lo[0]) / n[0], dy = (hi[1] - lo[1]) / n[1];
    C r = cmk(0, 0);
    for (size_t i = 0; i < n[0]; ++i)
        for (size_t j = 0; j < n[1]; ++j) {
            size_t idx = i * n[1] + j;
            double x = lo[0] + dx * i, y = lo[1] + dy * j;
            double ph = kx * x + ky * y;
            C e = cmk(cos(ph), -sin(ph));
            r = cadd(r, cmul(cscale(cadd(psi0[idx], dpsi[idx]), V[idx].re), e));
        }
    C t = cscale(r, (2 * mass) / (hbar * hbar) * dx * dy * (1. / 4));
    return cabs2(t) * (2 / (kPi * k0));
}

double orc_xsection3d(const size_t *n, const double *lo, const double *hi, double energy,
                      double mass, double hbar, const orc_cplx *psi0, const orc_cplx *dpsi,
                      const orc_cplx *V, double cosx, double cosy, double cosz,
                      orc_cplx *amp_out)
{
    /* ScatteringSolver3DImpl::ComputeXSection, src/ScatteringSolverImpl.cpp:318-361,
     * reduction op_Xsec3D (src/DeviceTemplate.h:1104-1134) */
    double nm = 1 / sqrt(cosx * cosx + cosy * cosy + cosz * cosz);
    cosx *= nm;
    cosy *= nm;
    cosz *= nm;
    double k0 = sqrt(2 * energy * mass) / hbar;
    double kx = k0 * cosx, ky = k0 * cosy, kz = k0 * cosz;
    double dx = (hi[0] - lo[0]) / n[0], dy = (hi[1] - lo[1]) / n[1], dz = (hi[2] - lo[2]) / n[2];
    C r = cmk(0, 0);
    for (size_t i = 0; i < n[0]; ++i)
        for (size_t j = 0; j < n[1]; ++j)
            for (size_t k = 0; k < n[2]; ++k) {
                size_t idx = (i * n[1] + j) * n[2] + k;
                double ph = (lo[0] + (double)i * dx) * kx + (lo[1] + (double)j * dy) * ky
                            + (lo[2] + (double)k * dz) * kz;
                C e = cmk(cos(-ph), sin(-ph));
                r = cadd(r, cmul(cscale(cadd(psi0[idx], dpsi[idx]), V[idx].re), e));
            }
    if (amp_out) *amp_out = r;
    C t = cscale(r, (2 * mass) / (hbar * hbar) * dx * dy * dz / (4 * kPi));
    return cabs2(t);
}

double orc_total_xsection2d(const size_t *n, const double *lo, const double *hi, double energy,
                            double mass, double hbar, const orc_cplx *psi0,
                            const orc_cplx *dpsi, const orc_cplx *V, long long nsample)
{
    /* src/ScatteringSolverImpl.cpp:199-212 */
    double t = 0;
    for (long long i = 0; i < nsample; ++i) {
        double th = i * 2 * kPi / nsample;
        t += orc_xsection2d(n, lo, hi, energy, mass, hbar, psi0, dpsi, V, cos(th), sin(th));
    }
    return t * 2 * kPi / nsample;
}

double orc_total_xsection3d(const size_t *n, const double *lo, const double *hi, double energy,
                            double mass, double hbar, const orc_cplx *psi0,
                            const orc_cplx *dpsi, const orc_cplx *V, long long npsi,
                            long long ntheta)
{
    /* src/ScatteringSolverImpl.cpp:363-379; theta indexed by j (the reference
     * indexes it by i at :369, a documented divergence) */
    double t = 0;
    for (long long j = 0; j < ntheta; ++j)
        for (long long i = 0; i < npsi; ++i) {
            double phi = i * 2 * kPi / npsi, th = j * kPi / ntheta;
            t += orc_xsection3d(n, lo, hi, energy, mass, hbar, psi0, dpsi, V,
                                sin(th) * cos(phi), sin(th) * sin(phi), cos(th), NULL)
                 * sin(th);
        }
    return t * 4 * kPi / (ntheta * npsi);
}
