/*
 * qsim_hip.h -- C ABI of libqsim_hip.so, the MI355X (gfx950) device behind
 * QuSim's Device / FourierTransform plug-in boundary.
 *
 * Plain C: pointers, sizes, doubles.  Every array argument is a DEVICE
 * pointer obtained from qsim_hip_malloc (or any HIP allocation on the current
 * device) unless the comment says "host".  Complex numbers are interleaved
 * (re, im), the layout of std::complex<double> / std::complex<float>.
 * Arrays are row-major, last axis contiguous: idx = i*ny + j (2D),
 * idx = (i*ny + j)*nz + k (3D)  (reference include/QuSim.h:47-61).
 *
 * Every function returns QSIM_OK (0) or an error code; the message of the
 * last error on the calling thread is qsim_hip_last_error().  A host adapter
 * turns a non-zero status into std::runtime_error, as the reference's
 * backends throw (src/CudaDevice.cu:31).
 *
 * Work is enqueued on `stream` (a hipStream_t; NULL = the null stream).
 * Reductions and element reads return their result to the host and
 * therefore synchronise that stream, as the reference's device ops are
 * synchronous (src/CudaDevice.cu:129-302).
 */
#ifndef QSIM_HIP_H
#define QSIM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSIM_OK 0
#define QSIM_ERR_INVALID 1      /* bad argument (size, pointer, option)       */
#define QSIM_ERR_HIP 2          /* HIP runtime error                          */
#define QSIM_ERR_NOMEM 3        /* device allocation failed                   */
#define QSIM_ERR_UNSUPPORTED 4  /* valid request the library does not support */

typedef struct { double re, im; } qsim_c128;
typedef struct { float re, im; } qsim_c64;
typedef void *qsim_stream_t;

/* ---------------------------------------------------------------- runtime */
const char *qsim_hip_last_error(void);
int qsim_hip_init(int device);                /* select device, set up library state */
int qsim_hip_device_info(char *name, size_t name_len, int *cu_count,
                         size_t *mem_free, size_t *mem_total);
int qsim_hip_synchronize(qsim_stream_t stream);
int qsim_hip_stream_create(qsim_stream_t *stream);
int qsim_hip_stream_destroy(qsim_stream_t stream);

/* Device::AllocMem / Free / MemToDevice / MemToHost (src/Device.h:53-80).
 * Copies are synchronous with respect to the host. */
int qsim_hip_malloc(void **ptr, size_t bytes);
int qsim_hip_free(void *ptr);
int qsim_hip_memcpy_h2d(void *dst, const void *host_src, size_t bytes);
int qsim_hip_memcpy_d2h(void *host_dst, const void *src, size_t bytes);
int qsim_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, qsim_stream_t stream);

/* Timing with HIP events on a stream (for measurement code). */
int qsim_hip_event_create(void **ev);
int qsim_hip_event_destroy(void *ev);
int qsim_hip_event_record(void *ev, qsim_stream_t stream);
int qsim_hip_event_elapsed_ms(void *start, void *stop, float *ms);
/* *done = 1 once the work before the event's record has completed, else 0
 * (non-blocking); event_synchronize waits for it. */
int qsim_hip_event_query(void *ev, int *done);
int qsim_hip_event_synchronize(void *ev);

/* ------------------------------------------------------- pointwise Device ops
 * Each replaces the DeviceTemplate<CudaExec> method of the same meaning
 * (src/Device.h:53-173, functors src/DeviceTemplate.h:531-1134). */
int qsim_hip_set_zero(qsim_c128 *a, size_t n, qsim_stream_t s);             /* SetZero  :1185 */
int qsim_hip_set_one(qsim_c128 *a, size_t n, qsim_stream_t s);              /* SetOne   :1192 */
int qsim_hip_copy(qsim_c128 *to, const qsim_c128 *from, size_t n, qsim_stream_t s); /* Copy :1313 */
int qsim_hip_copy_real(double *to, const double *from, size_t n, qsim_stream_t s);  /* Copy :1321 */
/* Scale(DComplex*, double) :1345 ; Scale(DComplex*, DComplex) :1353 */
int qsim_hip_scale(qsim_c128 *a, double scale, size_t n, qsim_stream_t s);
int qsim_hip_scale_c(qsim_c128 *a, qsim_c128 scale, size_t n, qsim_stream_t s);
/* Scale(to, from, double) :1370 */
int qsim_hip_scale_to(qsim_c128 *to, const qsim_c128 *from, double scale, size_t n,
                      qsim_stream_t s);
/* Mul(to, from2): to[i] = from2[i]*to[i]  :1458 ; Mul(to, a, b) :1414 */
int qsim_hip_mul(qsim_c128 *to, const qsim_c128 *from, size_t n, qsim_stream_t s);
int qsim_hip_mul3(qsim_c128 *to, const qsim_c128 *a, const qsim_c128 *b, size_t n,
                  qsim_stream_t s);
/* Exp(to, DReal* V, alpha): to = exp(alpha*V)  :1294 ; complex V :1312 */
int qsim_hip_exp_real(qsim_c128 *to, const double *V, qsim_c128 alpha, size_t n,
                      qsim_stream_t s);
int qsim_hip_exp_complex(qsim_c128 *to, const qsim_c128 *V, qsim_c128 alpha, size_t n,
                         qsim_stream_t s);
/* MulExp(psi, alpha, DReal* V): psi *= exp(alpha*V)  :1488 ; complex V :1479 */
int qsim_hip_mul_exp_real(qsim_c128 *psi, qsim_c128 alpha, const double *V, size_t n,
                          qsim_stream_t s);
int qsim_hip_mul_exp_complex(qsim_c128 *psi, qsim_c128 alpha, const qsim_c128 *V, size_t n,
                             qsim_stream_t s);
/* MulExpK1D/2D/3D: psi *= exp(alpha * sum_axis a_axis * fold(i)^2)  :1561-1589.
 * K1D takes the product alpha*a as one complex (reference signature). */
int qsim_hip_mul_exp_k1d(qsim_c128 *psi, qsim_c128 alpha, size_t n, qsim_stream_t s);
int qsim_hip_mul_exp_k2d(qsim_c128 *psi, qsim_c128 alpha, double ax, size_t nx, double ay,
                         size_t ny, qsim_stream_t s);
int qsim_hip_mul_exp_k3d(qsim_c128 *psi, qsim_c128 alpha, double ax, size_t nx, double ay,
                         size_t ny, double az, size_t nz, qsim_stream_t s);
/* MulExpIdx2: psi[i] *= exp(alpha * i^2), i < n  :1520 */
int qsim_hip_mul_exp_idx2(qsim_c128 *psi, qsim_c128 alpha, size_t n, qsim_stream_t s);
/* CopyReverseMinu(out, in, n): out[-m] = -in[m], m < n  :1303 (out points at the
 * LAST element written) */
int qsim_hip_copy_reverse_minus(qsim_c128 *out_last, const qsim_c128 *in, size_t n,
                                qsim_stream_t s);
/* Set / At single elements :1168-1180, :1155-1166 (At is synchronous) */
int qsim_hip_set_elem(qsim_c128 *a, size_t idx, qsim_c128 v, qsim_stream_t s);
int qsim_hip_get_elem(const qsim_c128 *a, size_t idx, qsim_c128 *host_out, qsim_stream_t s);

/* The remaining Device methods (src/Device.h:53-173) */
int qsim_hip_exp_i_real(qsim_c128 *to, const double *V, double alpha, size_t n,
                        qsim_stream_t s);                                      /* ExpI :1303 */
int qsim_hip_sub(qsim_c128 *to, const qsim_c128 *a, const qsim_c128 *b, size_t n,
                 qsim_stream_t s);                                             /* Sub :1383 */
int qsim_hip_mul_r(qsim_c128 *to, const qsim_c128 *a, const qsim_c128 *b, size_t n,
                   qsim_stream_t s);                                           /* MulR :1393 */
int qsim_hip_mul_minus_i(qsim_c128 *to, const qsim_c128 *a, const qsim_c128 *b, size_t n,
                         qsim_stream_t s);                                     /* MulMinusI :1403 */
int qsim_hip_mul_real_c(qsim_c128 *to, const double *a, const qsim_c128 *b, size_t n,
                        qsim_stream_t s);                                      /* Mul(DReal*) :1436 */
int qsim_hip_mul_real_alpha_c(qsim_c128 *to, const double *a, qsim_c128 alpha,
                              const qsim_c128 *b, size_t n, qsim_stream_t s); /* Mul(DReal*,a) :1446 */
int qsim_hip_mul_exp3(qsim_c128 *to, const qsim_c128 *a, qsim_c128 alpha, const qsim_c128 *b,
                      size_t n, qsim_stream_t s);                              /* MulExp(to,..) :1468 */
int qsim_hip_mul_exp_i_real(qsim_c128 *psi, double alpha, const double *V, size_t n,
                            qsim_stream_t s);                                  /* MulExpI :1498 */
int qsim_hip_scale_real(double *a, double scale, size_t n, qsim_stream_t s);   /* Scale(DReal*) */
int qsim_hip_scale_real_to(double *to, const double *from, double scale, size_t n,
                           qsim_stream_t s);
int qsim_hip_scale_c_to(qsim_c128 *to, const qsim_c128 *from, qsim_c128 scale, size_t n,
                        qsim_stream_t s);
int qsim_hip_set_elem_real(double *a, size_t idx, double v, qsim_stream_t s);
int qsim_hip_get_elem_real(const double *a, size_t idx, double *host_out, qsim_stream_t s);

/* Born-series pointwise ops */
int qsim_hip_g01d(qsim_c128 *psik, double E, double eps, double a, size_t n, qsim_stream_t s);
int qsim_hip_g02d(qsim_c128 *psik, double E, double eps, double a1, double a2, size_t nx,
                  size_t ny, qsim_stream_t s);                                  /* G02D :1622 */
int qsim_hip_g03d(qsim_c128 *psik, double E, double eps, double a1, double a2, double a3,
                  size_t nx, size_t ny, size_t nz, qsim_stream_t s);            /* G03D :1635 */
int qsim_hip_add_mul(qsim_c128 *psi, const qsim_c128 *psi0, const qsim_c128 *V, size_t n,
                     qsim_stream_t s);                                          /* AddMul :1667 */
int qsim_hip_add3(qsim_c128 *out, const qsim_c128 *dpsi, const qsim_c128 *psi0,
                  const qsim_c128 *V, double eps, size_t n, qsim_stream_t s);   /* Add3 :1689 */
int qsim_hip_linear_update(qsim_c128 *psi, const qsim_c128 *newpsi, double gamma, size_t n,
                           qsim_stream_t s);                                    /* :1700 */
int qsim_hip_vellekoop(qsim_c128 *psi, const qsim_c128 *newpsi, const qsim_c128 *V,
                       double slow, double eps, size_t n, qsim_stream_t s);     /* :1710 */
int qsim_hip_hao1(qsim_c128 *psi, const qsim_c128 *newpsi, const qsim_c128 *V, double slow,
                  double eps, size_t n, qsim_stream_t s);                       /* :1720 */
int qsim_hip_hao2(qsim_c128 *psi, const qsim_c128 *newpsi, const qsim_c128 *V, double slow,
                  double eps, size_t n, qsim_stream_t s);                       /* :1730 */

/* ------------------------------------------------------------ reductions
 * FP64 compensated (double-double) accumulation, deterministic order. */
int qsim_hip_norm2(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s);     /* :1222 */
int qsim_hip_norm2_real(const double *a, size_t n, double *out, qsim_stream_t s);   /* :1284 */
int qsim_hip_minus_norm2(const qsim_c128 *a, const qsim_c128 *b, size_t n, double *out,
                         qsim_stream_t s);                                          /* :1273 */
int qsim_hip_max_abs(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s);   /* Max :1241 */
int qsim_hip_min_abs(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s);   /* Min :1231 */
int qsim_hip_sum_real(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s);  /* :1251 */
int qsim_hip_sum_imag(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s);  /* :1262 */
int qsim_hip_abs2_idx(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s);  /* :1200 */
int qsim_hip_abs2_idx2(const qsim_c128 *a, size_t n, double *out, qsim_stream_t s); /* :1529 */
int qsim_hip_abs2_idx_real(const double *a, size_t n, double *out, qsim_stream_t s); /* :1210 */
int qsim_hip_abs2_k1d(const qsim_c128 *a, double ax, size_t n, double *out,
                      qsim_stream_t s);                                             /* :1539 */
int qsim_hip_abs2_k2d(const qsim_c128 *a, double ax, size_t nx, double ay, size_t ny,
                      double *out, qsim_stream_t s);                                /* :1550 */
int qsim_hip_abs2_k3d(const qsim_c128 *a, double ax, size_t nx, double ay, size_t ny,
                      double az, size_t nz, double *out, qsim_stream_t s);          /* :1561 */
int qsim_hip_abs2_mul_real(const qsim_c128 *a, const double *V, size_t n, double *out,
                           qsim_stream_t s);                                        /* :1601 */
int qsim_hip_abs2_mul_complex(const qsim_c128 *a, const qsim_c128 *V, size_t n,
                              qsim_c128 *out, qsim_stream_t s);                     /* :1591 */
/* Xsection3D: sum (psi0+psi) Re V exp(-i k.r)  :1646 */
int qsim_hip_xsection3d(const qsim_c128 *psi0, const qsim_c128 *psi, const qsim_c128 *V,
                        double kx, double ky, double kz, double x0, double y0, double z0,
                        double dx, double dy, double dz, size_t nx, size_t ny, size_t nz,
                        qsim_c128 *out, qsim_stream_t s);
/* 2D counterpart, replacing the host loop of src/ScatteringSolverImpl.cpp:186-191 */
int qsim_hip_xsection2d(const qsim_c128 *psi0, const qsim_c128 *psi, const qsim_c128 *V,
                        double kx, double ky, double x0, double y0, double dx, double dy,
                        size_t nx, size_t ny, qsim_c128 *out, qsim_stream_t s);

/* ------------------------------------------------------------------ FFT
 * FourierTransform{1,2,3}D::Create / Transform (src/FourierTransform.h:12-36):
 * unnormalised, forward sign -, inverse sign +, out of place (in == out is
 * also allowed).  dims outer -> inner.  precision 64: complex128 data;
 * precision 32: complex64 data (twiddles built in FP64, then rounded). */
typedef struct qsim_fft_plan qsim_fft_plan;
int qsim_hip_fft_plan_create(qsim_fft_plan **plan, int rank, const size_t *dims, int inverse,
                             int precision);
int qsim_hip_fft_exec(qsim_fft_plan *plan, const void *in, void *out, qsim_stream_t s);
int qsim_hip_fft_plan_destroy(qsim_fft_plan *plan);

/* ------------------------------------------------- fused split-operator job
 * SplittingMethod1D/2D::UpdatePsi (src/SplittingUtils.h:14-64,
 * src/SplittingMethod1D.cpp, src/SplittingMethod2D.cpp) as fused sweeps:
 * each FFT is fused with the phase multiply that follows it, so a 2D step
 * is two read+write passes over psi. */
typedef struct {
    int dim;              /* 1 or 2                                          */
    size_t nx, ny;        /* ny = 1 for 1D                                   */
    double x0, x1, y0, y1;
    double dt_re, dt_im;  /* complex dt as the reference (include/QuSim.h:558) */
    double mass, hbar;
    int method;           /* 2 = SplittingMethodO2, 4 = SplittingMethodO4    */
    int boundary;         /* 0 = Period, 1 = InfiniteWall (1D)               */
    int precision;        /* 64 = complex128; 32 = FP32-mixed                */
} qsim_split_desc;

typedef struct qsim_split_job qsim_split_job;
/* V_re: device, n doubles (the reference's real potential fV).  V_im: device,
 * n doubles or NULL (complex-potential extension; NULL reproduces the
 * reference evolver, which keeps Re V only, src/EvolverImpl2D.cpp:55).
 * psi: device complex128 state, updated in place by qsim_hip_split_steps. */
int qsim_hip_split_create(qsim_split_job **job, const qsim_split_desc *d, const double *V_re,
                          const double *V_im, qsim_c128 *psi);
/* nsteps split steps (QuUpdatePsi with fBatchSize = nsteps); psi is in real
 * space and up to date when the enqueued work completes. */
int qsim_hip_split_steps(qsim_split_job *job, size_t nsteps, qsim_stream_t s);
/* As qsim_hip_split_steps, timing every kernel launch with HIP events on
 * `stream`: kernel_ms[k] / kernel_launches[k] (arrays of 5) for k = 0 row
 * sweeps between steps (IFFT_y, E_V, FFT_y), 1 column sweeps, 2 the
 * persistent 1D kernel, 3 the first row sweep of the call (E_V, FFT_y),
 * 4 the last (IFFT_y, E_V).  Synchronous.  Fused jobs only. */
int qsim_hip_split_steps_timed(qsim_split_job *job, size_t nsteps, double *kernel_ms,
                               int *kernel_launches, qsim_stream_t s);
/* 2D fused jobs replay each step count's sweep sequence from a captured HIP
 * graph (off by default: replay and direct launches measured the same on
 * MI355X for a single job); enable = 1 turns it on.  At most 8 step counts
 * are kept per job (the oldest capture is dropped first); results are bitwise
 * those of direct launches. */
int qsim_hip_split_use_graph(qsim_split_job *job, int enable);
/* Launch variants of the 2D sweep kernels (radix set, lines per workgroup):
 * 0..7 for axis lengths 1024 and 2048 (6 and 7 run four radix passes with
 * 8 elements per thread), only 0 otherwise.  A new job starts with a fixed
 * pair measured fastest on MI355X for its lengths (FP64 1024: 4/4, FP32
 * 1024: 7/7, FP64 2048: 7/5, FP32 2048: 7/3). */
int qsim_hip_split_set_variants(qsim_split_job *job, int row_variant, int col_variant);
/* x-split factor R of a fused 2D job (1, 2 or 4): the x transform is taken
 * as R-point transforms across the rows {i2 + k nx/R} inside the row sweep
 * (lane exchanges, with the four-step twiddles) and transforms of nx / R
 * points in the column sweep, which then has R * ny lines.  The same
 * mathematics in a different order: results agree with R = 1 to rounding.
 * Setting R resets the column variant to the default of the new column
 * length.  QSIM_ERR_UNSUPPORTED when nx / R has no fused kernel or a group
 * of R rows does not fit one workgroup. */
int qsim_hip_split_set_xsplit(qsim_split_job *job, int xsplit);
/* The job's current launch schedule (after autotuning, for example). */
int qsim_hip_split_get_schedule(const qsim_split_job *job, int *row_variant, int *col_variant,
                                int *xsplit);
/* Autotuner: for each x-split factor R (1, 2, 4) the job can take, times
 * every row variant, then every column variant, each in reps alternating
 * (row, column) sweep pairs on a copy of the state; keeps the fastest
 * (R, row, column) schedule on the job (qsim_hip_split_get_schedule reads
 * it) and reports its variants.  *row_us / *col_us: microseconds per pair at
 * the end of the row and of the column stage.  Synchronous. */
int qsim_hip_split_autotune(qsim_split_job *job, int reps, int *row_variant, int *col_variant,
                            double *row_us, double *col_us, qsim_stream_t s);
/* The same search with all njobs same-shape jobs in every launch, as
 * qsim_hip_split_steps_batched and qsim_hip_split_batch_steps run them (more
 * workgroups per launch can favour other variants); sets the chosen pair on
 * every job.  Synchronous.  No counterpart in the reference (its solvers run
 * one grid at a time). */
int qsim_hip_split_autotune_batched(qsim_split_job *const *jobs, int njobs, int reps,
                                    int *row_variant, int *col_variant, double *row_us,
                                    double *col_us, qsim_stream_t s);
/* Re-read psi after the caller changed it (e.g. renormalised it). */
int qsim_hip_split_load_psi(qsim_split_job *job, qsim_stream_t s);
/* sum |psi|^2 over the job's working state (compensated FP64), synchronous */
int qsim_hip_split_norm2(qsim_split_job *job, double *out, qsim_stream_t s);
int qsim_hip_split_destroy(qsim_split_job *job);

/* Many same-shape split jobs advanced in one launch per sweep. */
int qsim_hip_split_steps_batched(qsim_split_job **jobs, int njobs, size_t nsteps,
                                 qsim_stream_t s);

/* No counterpart in the reference.  qsim_hip_split_steps_batched for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * request loops): the same sweeps and bitwise the same results, with the
 * descriptor table in a static per-GPU device array instead of a hipMalloc'd
 * one, written by a small kernel on `s` (no host staging).  It calls no
 * hipMalloc/hipFree and returns once the work is enqueued.
 * Limits: at most 512 jobs per call (QSIM_ERR_INVALID beyond); fused jobs only
 * (power-of-two axes, periodic boundary; QSIM_ERR_UNSUPPORTED otherwise).
 * Every call on one GPU shares that GPU's table, which keeps the last call's
 * descriptors: calls on one stream are ordered, calls that overlap on one GPU
 * from different streams or threads race. */
int qsim_hip_split_steps_batched_noalloc(qsim_split_job *const *jobs, int njobs, size_t nsteps,
                                         qsim_stream_t s);

/* Reusable asynchronous batch (request loops): one launch per sweep for up
 * to `capacity` same-shape fused jobs that share method, precision and
 * launch variants.  batch_steps returns once the work is enqueued; the
 * batch's descriptor table is reused by the next call, so one batch object
 * serves one stream at a time.  Each job's result is bitwise the same as
 * qsim_hip_split_steps on that job alone with the same nsteps. */
typedef struct qsim_split_batch qsim_split_batch;
int qsim_hip_split_batch_create(qsim_split_batch **batch, int capacity);
int qsim_hip_split_batch_steps(qsim_split_batch *batch, qsim_split_job *const *jobs, int njobs,
                               size_t nsteps, qsim_stream_t s);
int qsim_hip_split_batch_destroy(qsim_split_batch *batch);

/* ------------------------------------------------ multi-size job groups
 * No counterpart in the reference (its solvers advance one grid at a time,
 * QuUpdatePsi in src/SplittingUtils.h:14-64, Batch(n) steps per call at
 * src/EvolverImpl.cpp:19).  A group advances fused 2D jobs of different grid
 * sizes (axes powers of two from 64 to 2048; one precision and method per
 * group; default launch schedule) together: each advance of `passes` passes
 * is one launch per sweep for all member jobs, from a fixed grid that loops
 * over a device work list, optionally replayed from a captured HIP graph.
 * A job added with nsteps starts at the next advance and runs the sweep
 * sequence of qsim_hip_split_steps(job, nsteps) -- bitwise the same result --
 * spread over as many advances as it needs (nph passes per step: 1 for O2, 3
 * for O4, plus one closing row pass); jobs join and leave between advances
 * without a new capture.  psi of a job is final once status reports it
 * finished and the enqueued work has completed; a job removed before it
 * finishes is left mid-step (its psi is not meaningful).  One stream per
 * group: advances read tables that the next add/remove rewrites in stream
 * order. */
typedef struct qsim_split_group qsim_split_group;
int qsim_hip_split_group_create(qsim_split_group **group, int capacity, int precision,
                                int method);
int qsim_hip_split_group_add(qsim_split_group *group, qsim_split_job *job, size_t nsteps);
int qsim_hip_split_group_remove(qsim_split_group *group, qsim_split_job *job);
int qsim_hip_split_group_advance(qsim_split_group *group, int passes, int use_graph,
                                 qsim_stream_t s);
int qsim_hip_split_group_status(const qsim_split_group *group, const qsim_split_job *job,
                                int *finished, long long *passes_left);
/* Member jobs, shape classes, cached captures, and the advances that ran as
 * direct launches because the runtime refused or invalidated their capture
 * (graph advances fall back to direct launches, bitwise the same results).
 * Any out-pointer may be NULL. */
int qsim_hip_split_group_info(const qsim_split_group *group, int *jobs, int *shapes,
                              int *graphs, long long *capture_failures);
int qsim_hip_split_group_destroy(qsim_split_group *group);

/* ------------------------------------------------------ fused Born job
 * QuPerturbation2DImpl/3DImpl (src/Perturbation2DImpl.cpp,
 * src/Perturbation3DImpl.cpp) with the kernels of src/Perburbation.h. */
typedef struct {
    int dim;                 /* 2 or 3                                        */
    size_t n[3];
    double lo[3], hi[3];
    double direction[3];     /* incident direction (normalised internally)     */
    double mass, hbar;
    int preconditional;      /* 0 naive BornSeries, 1 PreconditionalBornSeries */
    int preconditioner;      /* 0 Identity, 1 Vellekoop, 2 Hao1, 3 Hao2        */
    double slow;             /* relaxation (reference "slow", default 1)       */
    double epsilon;          /* user epsilon (raised to the minimum when preconditional) */
    int precision;           /* 64 or 32 (FP32 iterates, FP64 residual)        */
} qsim_born_desc;

typedef struct qsim_born_job qsim_born_job;
/* V: device complex128 physical potential (n points).  energies: host array
 * of nE energies; all energies share V and the grid and iterate together. */
int qsim_hip_born_create(qsim_born_job **job, const qsim_born_desc *d, const qsim_c128 *V,
                         const double *energies, int nE);
/* Up to max_iter iterations per energy; an energy stops once its relative
 * update norm |dpsi_new - dpsi_old| / |dpsi_new| drops below tol (tol <= 0:
 * run exactly max_iter, the reference's `order` semantics).  iters[nE] and
 * residual[nE] are host outputs (may be NULL). */
int qsim_hip_born_iterate(qsim_born_job *job, int max_iter, double tol, int *iters,
                          double *residual, qsim_stream_t s);
/* Copy energy e's scattered wave (complex128) to a device buffer. */
int qsim_hip_born_get_dpsi(qsim_born_job *job, int e, qsim_c128 *out, qsim_stream_t s);
/* Potential after the absorbing layer and the epsilon in use, energy e. */
int qsim_hip_born_get_veff(qsim_born_job *job, int e, qsim_c128 *out, double *eps,
                           qsim_stream_t s);
/* Differential cross section toward (cosx, cosy, cosz) for energy e
 * (src/ScatteringSolverImpl.cpp:168-196, :318-361); amp = the raw sum r. */
int qsim_hip_born_xsection(qsim_born_job *job, int e, double cosx, double cosy, double cosz,
                           double *dsigma, qsim_c128 *amp, qsim_stream_t s);
int qsim_hip_born_destroy(qsim_born_job *job);

/* ------------------------------------------------ input generation on the GPU
 * Replace the reference's per-point host evaluation of the initial psi and
 * the potential (std::function callbacks, src/EvolverImpl1D.cpp:40-70,
 * src/EvolverImpl2D.cpp:40-99, src/Perturbation3DImpl.cpp:28-50) for the
 * parameterised families of the benchmark workloads.  dim = 1..3; n, lo, hi
 * (host, dim entries) give the grid x_a[i] = lo[a] + (hi[a]-lo[a])/n[a]*i.
 *
 * Potential: V = sum_b h_b exp(-|x - c_b|^2 / (2 w_b^2))
 *               - i * absorb_strength * ((absorb_width - d)/absorb_width)^2
 * for d < absorb_width (d = distance to the nearest box edge; absorb_width
 * <= 0 disables the layer).  bumps (host): count x (h, w, c_0..c_{dim-1}),
 * count <= 64.  Outputs: re and/or im (n doubles) and/or cv (complex128);
 * NULL outputs are skipped. */
int qsim_hip_gen_potential(int dim, const size_t *n, const double *lo, const double *hi,
                           const double *bumps, int count, double absorb_width,
                           double absorb_strength, double *re, double *im, qsim_c128 *cv,
                           qsim_stream_t s);
/* Packet psi = exp(i k.x - |x - center|^2 / (4 sigma^2)); normalise != 0
 * scales it to sum |psi|^2 dV = 1 (compensated norm; synchronises s). */
int qsim_hip_gen_packet(int dim, const size_t *n, const double *lo, const double *hi,
                        const double *center, const double *k, double sigma, int normalise,
                        qsim_c128 *psi, qsim_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
