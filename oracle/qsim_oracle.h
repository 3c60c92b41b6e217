/*
 * qsim_oracle.h -- CPU oracle for the split-operator / Born-series path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C, FP64 restatement of
 * the reference QuSim CPU path (CpuSeq device + KISS FFT).  It is the checker
 * for the HIP product path: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * Each function cites the reference file:line it restates.  Pinning (what
 * checks this restatement independently) is listed in the header comment of
 * qsim_oracle.c.
 *
 * Complex numbers are interleaved (re, im) doubles, the layout of
 * std::complex<double>.  Arrays are row-major with the last axis contiguous:
 * idx = i*ny + j (2D), idx = (i*ny + j)*nz + k (3D), reference
 * include/QuSim.h:47-61.
 */
#ifndef QSIM_ORACLE_H
#define QSIM_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_cplx;

/* Unnormalised DFT, forward sign -, inverse sign +, reference
 * src/FourierTransform.h:16-19.  rank 1..3; dims outer->inner.  Axis order
 * follows src/FourierTransform.cpp:248-284 (2D: inner axis first) and
 * :404-447 (3D: z, then y, then x).  in may equal out. */
int orc_fft(int rank, const size_t *dims, int inverse,
            const orc_cplx *in, orc_cplx *out, int threads);

/* O(n^2) DFT in long double, for pinning the FFT on small sizes. */
int orc_dft_bruteforce(size_t n, int inverse, const orc_cplx *in, orc_cplx *out);

/* Plain (reference-order) sum of |a_i|^2, src/DeviceTemplate.h:1222-1229. */
double orc_norm2(const orc_cplx *a, size_t n);

/* Split-operator evolution.
 *   method   : 2 = SplittingMethodO2, 4 = SplittingMethodO4 (src/SplittingUtils.h:14-64)
 *   boundary : 0 = Period, 1 = InfiniteWall (1D only, src/SplittingMethod1D.cpp:62-81)
 *   V        : complex potential; when complex_v == 0 only V.re is used, which is
 *              exactly the reference evolver (src/EvolverImpl1D.cpp:46,
 *              src/EvolverImpl2D.cpp:55 keep the real part only)
 *   psi      : in/out, already normalised by the caller if wanted
 *   nsteps   : number of step() calls; each call does `batch` split steps
 *   fnes     : force normalisation after each step() call (src/EvolverImpl.cpp:60-63)
 */
int orc_split1d_run(size_t n, double x0, double x1, double dt_re, double dt_im,
                    double mass, double hbar, int method, int boundary,
                    const orc_cplx *V, int complex_v, orc_cplx *psi,
                    size_t nsteps, size_t batch, int fnes, int threads);

int orc_split2d_run(size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                    double dt_re, double dt_im, double mass, double hbar, int method,
                    const orc_cplx *V, int complex_v, orc_cplx *psi,
                    size_t nsteps, size_t batch, int fnes, int threads);

/* Time the 2D O2 step loop only (no set-up), returns seconds for nsteps
 * steps.  threads: the pointwise ops (Mul, Scale); fft_threads: the lines of
 * each 2D transform.  The reference's CpuPar threads only the former (its
 * KISS FFT is serial): threads = cores, fft_threads = 1. */
double orc_split2d_time(size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                        double dt, double mass, double hbar,
                        const orc_cplx *V, int complex_v, orc_cplx *psi,
                        size_t nsteps, int threads, int fft_threads);

/* Observables of the 1D evolver, src/EvolverImpl1D.cpp:82-101. */
double orc_norm2_range_1d(const orc_cplx *psi, size_t n, double x0, double x1,
                          double a, double b);

/* Evolver observables of psi on a 1D (ny ignored) or 2D grid with a real
 * potential V: out[0] KinEn (src/SplittingMethod1D.cpp:94-128,
 * src/SplittingMethod2D.cpp:124-141; boundary 1 = infinite wall, 1D only),
 * out[1] PotEn (src/EvolverImpl.cpp:45-49), out[2] Norm2 times dv
 * (src/EvolverImpl.cpp:51-54), out[3] Xavg (1D, src/EvolverImpl1D.cpp:82-87). */
int orc_observables(int dim, size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                    double mass, double hbar, int boundary, const orc_cplx *psi,
                    const double *V, double *out);

/* Born series.
 *   dim          : 2 or 3
 *   n[dim], lo[dim], hi[dim] : grid (x_i = lo + i*(hi-lo)/n)
 *   V            : physical complex potential (before the absorbing layer)
 *   preconditional, preconditioner (0 Identity, 1 Vellekoop, 2 Hao1, 3 Hao2),
 *   slow, order  : src/PerturbationOptions.cpp:6-43
 *   dpsi         : in/out scattered wave (zero on first call)
 *   v_eff        : out (may be NULL): V after the absorbing layer
 *   eps_inout    : in user epsilon, out the epsilon used (src/Perburbation.h:329-344)
 *   delta_norm   : out: sqrt(|dpsi-last|^2/|dpsi|^2) after the calls
 *                  (src/Perturbation2DImpl.cpp:80-82)
 *   ncalls       : number of Compute() calls, each doing `order` iterations
 */
int orc_born_run(int dim, const size_t *n, const double *lo, const double *hi,
                 double energy, const double *direction, double mass, double hbar,
                 int preconditional, int preconditioner, double slow, int order,
                 const orc_cplx *V, orc_cplx *dpsi, orc_cplx *v_eff,
                 double *eps_inout, double *delta_norm, int ncalls, int threads);

/* Seconds of `iters` Born iterations (the Compute() loop only, set-up and
 * the absorbing layer excluded) from dpsi = 0; threads / fft_threads as for
 * orc_split2d_time.  Returns a negative value on bad arguments. */
double orc_born_time(int dim, const size_t *n, const double *lo, const double *hi,
                     double energy, const double *direction, double mass, double hbar,
                     int preconditional, int preconditioner, double slow, int iters,
                     const orc_cplx *V, int threads, int fft_threads);

/* Plane wave psi0 = exp(i k.r), src/ScatteringSolverImpl.cpp:272-314. */
int orc_plane_wave(int dim, const size_t *n, const double *lo, const double *hi,
                   double energy, const double *direction, double mass, double hbar,
                   orc_cplx *psi0);

/* Cross sections, src/ScatteringSolverImpl.cpp:168-212 (2D) and :318-379 (3D).
 * V is the potential used by the solver (only Re V enters). */
double orc_xsection2d(const size_t *n, const double *lo, const double *hi,
                      double energy, double mass, double hbar,
                      const orc_cplx *psi0, const orc_cplx *dpsi, const orc_cplx *V,
                      double cosx, double cosy);
double orc_xsection3d(const size_t *n, const double *lo, const double *hi,
                      double energy, double mass, double hbar,
                      const orc_cplx *psi0, const orc_cplx *dpsi, const orc_cplx *V,
                      double cosx, double cosy, double cosz, orc_cplx *amp_out);
double orc_total_xsection2d(const size_t *n, const double *lo, const double *hi,
                            double energy, double mass, double hbar,
                            const orc_cplx *psi0, const orc_cplx *dpsi,
                            const orc_cplx *V, long long nsample);
double orc_total_xsection3d(const size_t *n, const double *lo, const double *hi,
                            double energy, double mass, double hbar,
                            const orc_cplx *psi0, const orc_cplx *dpsi,
                            const orc_cplx *V, long long npsi, long long ntheta);

#ifdef __cplusplus
}
#endif
#endif
