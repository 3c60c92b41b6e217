// qsim_ops.hip -- the Device interface's pointwise ops and reductions on
// gfx950.  Each op restates the DeviceTemplate functor cited beside it
// (src/DeviceTemplate.h); reductions accumulate in double-double so the FP64
// sums are compensated and independent of the grid shape's rounding.
#include <cmath>
#include <cstring>

#include "qsim_common.h"
#include "qsim_reduce.h"

using namespace qsim;

namespace {

using namespace qsim::red;

inline const c128 *C(const qsim_c128 *p) { return reinterpret_cast<const c128 *>(p); }
inline c128 *C(qsim_c128 *p) { return reinterpret_cast<c128 *>(p); }

template <int KIND>
int precond_update(qsim_c128 *psi_, const qsim_c128 *np_, const qsim_c128 *V_,
                          double slow, double eps, size_t n, qsim_stream_t s)
{
    QSIM_CHECK((psi_ && np_ && V_) || n == 0, "null device pointer");
    c128 *psi = C(psi_);
    const c128 *np = C(np_), *V = C(V_);
    return launch_map(n, s, [=] __device__(size_t i) {
        c128 g = born_gamma<KIND>(V[i], slow, eps);
        psi[i] = mk(1. - g.x, -g.y) * psi[i] + g * np[i];
    });
}

}  // namespace

#define NEED(p) QSIM_CHECK((p) != nullptr || n == 0, "null device pointer")

extern "C" {

// ------------------------------------------------------------- pointwise
int qsim_hip_set_zero(qsim_c128 *a_, size_t n, qsim_stream_t s)
{
    NEED(a_);
    c128 *a = C(a_);
    return launch_map(n, s, [=] __device__(size_t i) { a[i] = mk(0.0, 0.0); });
}

int qsim_hip_set_one(qsim_c128 *a_, size_t n, qsim_stream_t s)
{
    NEED(a_);
    c128 *a = C(a_);
    return launch_map(n, s, [=] __device__(size_t i) { a[i] = mk(1.0, 0.0); });
}

int qsim_hip_copy(qsim_c128 *to_, const qsim_c128 *from_, size_t n, qsim_stream_t s)
{
    NEED(to_);
    NEED(from_);
    c128 *to = C(to_);
    const c128 *from = C(from_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = from[i]; });
}

int qsim_hip_copy_real(double *to, const double *from, size_t n, qsim_stream_t s)
{
    NEED(to);
    NEED(from);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = from[i]; });
}

// op_Scale: alpha * o1 (src/DeviceTemplate.h:630-641)
int qsim_hip_scale(qsim_c128 *a_, double sc, size_t n, qsim_stream_t s)
{
    NEED(a_);
    c128 *a = C(a_);
    return launch_map(n, s, [=] __device__(size_t i) { a[i] = sc * a[i]; });
}

int qsim_hip_scale_c(qsim_c128 *a_, qsim_c128 sc_, size_t n, qsim_stream_t s)
{
    NEED(a_);
    c128 *a = C(a_);
    c128 sc = to_c(sc_);
    return launch_map(n, s, [=] __device__(size_t i) { a[i] = sc * a[i]; });
}

int qsim_hip_scale_to(qsim_c128 *to_, const qsim_c128 *from_, double sc, size_t n,
                      qsim_stream_t s)
{
    NEED(to_);
    NEED(from_);
    c128 *to = C(to_);
    const c128 *from = C(from_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = sc * from[i]; });
}

// Mul(to, from2): op_Mul(from2[i], to[i]) (src/DeviceTemplate.h:1458-1466)
int qsim_hip_mul(qsim_c128 *to_, const qsim_c128 *from_, size_t n, qsim_stream_t s)
{
    NEED(to_);
    NEED(from_);
    c128 *to = C(to_);
    const c128 *from = C(from_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = from[i] * to[i]; });
}

int qsim_hip_mul3(qsim_c128 *to_, const qsim_c128 *a_, const qsim_c128 *b_, size_t n,
                  qsim_stream_t s)
{
    NEED(to_);
    NEED(a_);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *a = C(a_), *b = C(b_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = a[i] * b[i]; });
}

// op_ExpMul: exp(alpha * o1) (src/DeviceTemplate.h:604-615)
int qsim_hip_exp_real(qsim_c128 *to_, const double *V, qsim_c128 alpha_, size_t n,
                      qsim_stream_t s)
{
    NEED(to_);
    NEED(V);
    c128 *to = C(to_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = cexp(alpha * V[i]); });
}

int qsim_hip_exp_complex(qsim_c128 *to_, const qsim_c128 *V_, qsim_c128 alpha_, size_t n,
                         qsim_stream_t s)
{
    NEED(to_);
    NEED(V_);
    c128 *to = C(to_);
    const c128 *V = C(V_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = cexp(alpha * V[i]); });
}

// op_MulExpMul: o1 * exp(alpha * o2) (src/DeviceTemplate.h:617-628)
int qsim_hip_mul_exp_real(qsim_c128 *psi_, qsim_c128 alpha_, const double *V, size_t n,
                          qsim_stream_t s)
{
    NEED(psi_);
    NEED(V);
    c128 *psi = C(psi_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s,
                      [=] __device__(size_t i) { psi[i] = psi[i] * cexp(alpha * V[i]); });
}

int qsim_hip_mul_exp_complex(qsim_c128 *psi_, qsim_c128 alpha_, const qsim_c128 *V_,
                             size_t n, qsim_stream_t s)
{
    NEED(psi_);
    NEED(V_);
    c128 *psi = C(psi_);
    const c128 *V = C(V_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s,
                      [=] __device__(size_t i) { psi[i] = psi[i] * cexp(alpha * V[i]); });
}

// op_MulExpK1D (src/DeviceTemplate.h:754-769)
int qsim_hip_mul_exp_k1d(qsim_c128 *psi_, qsim_c128 ax_, size_t n, qsim_stream_t s)
{
    NEED(psi_);
    c128 *psi = C(psi_);
    c128 ax = to_c(ax_);
    return launch_map(n, s, [=] __device__(size_t i) {
        double k = (double)fold_half(i, n);
        psi[i] = psi[i] * cexp(ax * (k * k));
    });
}

// op_MulExpK2D (src/DeviceTemplate.h:794-819)
int qsim_hip_mul_exp_k2d(qsim_c128 *psi_, qsim_c128 alpha_, double ax, size_t nx, double ay,
                         size_t ny, qsim_stream_t s)
{
    size_t n = nx * ny;
    NEED(psi_);
    c128 *psi = C(psi_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t idx) {
        double i = (double)fold_half(idx / ny, nx), j = (double)fold_half(idx % ny, ny);
        psi[idx] = psi[idx] * cexp(alpha * (i * i * ax + j * j * ay));
    });
}

// op_MulExpK3D (src/DeviceTemplate.h:852-879).  The reference drops alpha at
// :876; this applies exp(alpha * T) as its 1D/2D siblings do.
int qsim_hip_mul_exp_k3d(qsim_c128 *psi_, qsim_c128 alpha_, double ax, size_t nx, double ay,
                         size_t ny, double az, size_t nz, qsim_stream_t s)
{
    size_t n = nx * ny * nz;
    NEED(psi_);
    c128 *psi = C(psi_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t idx) {
        size_t k = idx % nz, r = idx / nz;
        double fk = (double)fold_half(k, nz), fj = (double)fold_half(r % ny, ny),
               fi = (double)fold_half(r / ny, nx);
        psi[idx] = psi[idx] * cexp(alpha * (fi * fi * ax + fj * fj * ay + fk * fk * az));
    });
}

// op_MulExpSquare (src/DeviceTemplate.h:979-990)
int qsim_hip_mul_exp_idx2(qsim_c128 *psi_, qsim_c128 alpha_, size_t n, qsim_stream_t s)
{
    NEED(psi_);
    c128 *psi = C(psi_);
    c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t i) {
        double k = (double)i;
        psi[i] = psi[i] * cexp(alpha * (k * k));
    });
}

// CopyReverseMinu (src/DeviceTemplate.h:1303-1310): out[-m] = -in[m]
int qsim_hip_copy_reverse_minus(qsim_c128 *out_, const qsim_c128 *in_, size_t n,
                                qsim_stream_t s)
{
    NEED(out_);
    NEED(in_);
    c128 *out = C(out_);
    const c128 *in = C(in_);
    return launch_map(n, s, [=] __device__(size_t m) {
        c128 v = in[m];
        *(out - m) = mk(-v.x, -v.y);
    });
}

int qsim_hip_set_elem(qsim_c128 *a, size_t idx, qsim_c128 v, qsim_stream_t s)
{
    QSIM_CHECK(a, "null device pointer");
    // the value travels by kernel argument, so no host buffer has to outlive the call
    c128 *p = C(a) + idx;
    c128 val = to_c(v);
    return launch_map(1, s, [=] __device__(size_t) { *p = val; });
}

int qsim_hip_get_elem(const qsim_c128 *a, size_t idx, qsim_c128 *out, qsim_stream_t s)
{
    QSIM_CHECK(a && out, "null pointer");
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_HIP(hipStreamSynchronize(as_stream(s)));
    QSIM_HIP(hipMemcpy(out, a + idx, sizeof(qsim_c128), hipMemcpyDeviceToHost));
    return QSIM_OK;
}

// ------------------------------------------------ remaining Device methods
// op_ExpI: exp(i * alpha * V) (src/DeviceTemplate.h:570-580)
int qsim_hip_exp_i_real(qsim_c128 *to_, const double *V, double alpha, size_t n, qsim_stream_t s)
{
    NEED(to_);
    NEED(V);
    c128 *to = C(to_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = cexp(mk(0.0, V[i] * alpha)); });
}

// op_Sub (src/DeviceTemplate.h:688-694)
int qsim_hip_sub(qsim_c128 *to_, const qsim_c128 *a_, const qsim_c128 *b_, size_t n,
                 qsim_stream_t s)
{
    NEED(to_);
    NEED(a_);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *a = C(a_), *b = C(b_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = a[i] - b[i]; });
}

// op_MulR: o1 * Re(o2) (src/DeviceTemplate.h:709-715)
int qsim_hip_mul_r(qsim_c128 *to_, const qsim_c128 *a_, const qsim_c128 *b_, size_t n,
                   qsim_stream_t s)
{
    NEED(to_);
    NEED(a_);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *a = C(a_), *b = C(b_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = a[i] * b[i].x; });
}

// op_MulMinusI: o1 * (0, -Im o2) (src/DeviceTemplate.h:717-723)
int qsim_hip_mul_minus_i(qsim_c128 *to_, const qsim_c128 *a_, const qsim_c128 *b_, size_t n,
                         qsim_stream_t s)
{
    NEED(to_);
    NEED(a_);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *a = C(a_), *b = C(b_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = a[i] * mk(0.0, -b[i].y); });
}

// Mul(to, DReal* from1, from2): from1 * from2 (src/DeviceTemplate.h:1436-1444)
int qsim_hip_mul_real_c(qsim_c128 *to_, const double *a, const qsim_c128 *b_, size_t n,
                        qsim_stream_t s)
{
    NEED(to_);
    NEED(a);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *b = C(b_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = a[i] * b[i]; });
}

// op_MulMul: o1 * alpha * o2 (src/DeviceTemplate.h:725-734)
int qsim_hip_mul_real_alpha_c(qsim_c128 *to_, const double *a, qsim_c128 alpha_,
                              const qsim_c128 *b_, size_t n, qsim_stream_t s)
{
    NEED(to_);
    NEED(a);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *b = C(b_);
    const c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = (a[i] * alpha) * b[i]; });
}

// MulExp(to, from1, alpha, from2): from1 * exp(alpha * from2) (src/DeviceTemplate.h:1468-1476)
int qsim_hip_mul_exp3(qsim_c128 *to_, const qsim_c128 *a_, qsim_c128 alpha_, const qsim_c128 *b_,
                      size_t n, qsim_stream_t s)
{
    NEED(to_);
    NEED(a_);
    NEED(b_);
    c128 *to = C(to_);
    const c128 *a = C(a_), *b = C(b_);
    const c128 alpha = to_c(alpha_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = a[i] * cexp(alpha * b[i]); });
}

// op_MulExpI: o1 * IMul(alpha * o2) -- a multiply by i alpha V, not an
// exponential, exactly as the reference functor (src/DeviceTemplate.h:589-600)
int qsim_hip_mul_exp_i_real(qsim_c128 *psi_, double alpha, const double *V, size_t n,
                            qsim_stream_t s)
{
    NEED(psi_);
    NEED(V);
    c128 *psi = C(psi_);
    return launch_map(n, s, [=] __device__(size_t i) { psi[i] = psi[i] * mk(0.0, alpha * V[i]); });
}

int qsim_hip_scale_real(double *a, double sc, size_t n, qsim_stream_t s)
{
    NEED(a);
    return launch_map(n, s, [=] __device__(size_t i) { a[i] = sc * a[i]; });
}

int qsim_hip_scale_real_to(double *to, const double *from, double sc, size_t n, qsim_stream_t s)
{
    NEED(to);
    NEED(from);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = sc * from[i]; });
}

int qsim_hip_scale_c_to(qsim_c128 *to_, const qsim_c128 *from_, qsim_c128 sc_, size_t n,
                        qsim_stream_t s)
{
    NEED(to_);
    NEED(from_);
    c128 *to = C(to_);
    const c128 *from = C(from_);
    const c128 sc = to_c(sc_);
    return launch_map(n, s, [=] __device__(size_t i) { to[i] = sc * from[i]; });
}

int qsim_hip_set_elem_real(double *a, size_t idx, double v, qsim_stream_t s)
{
    QSIM_CHECK(a, "null device pointer");
    double *p = a + idx;
    return launch_map(1, s, [=] __device__(size_t) { *p = v; });
}

int qsim_hip_get_elem_real(const double *a, size_t idx, double *out, qsim_stream_t s)
{
    QSIM_CHECK(a && out, "null pointer");
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_HIP(hipStreamSynchronize(as_stream(s)));
    QSIM_HIP(hipMemcpy(out, a + idx, sizeof(double), hipMemcpyDeviceToHost));
    return QSIM_OK;
}

// ------------------------------------------------------------ Born ops
// op_G01D (src/DeviceTemplate.h:882-903)
int qsim_hip_g01d(qsim_c128 *psik_, double E, double eps, double a, size_t n,
                  qsim_stream_t s)
{
    NEED(psik_);
    c128 *psik = C(psik_);
    return launch_map(n, s, [=] __device__(size_t i) {
        double k = (double)fold_half(i, n);
        psik[i] = cdiv(psik[i], mk(E - k * k * a, eps));
    });
}

// op_G02D (src/DeviceTemplate.h:905-934)
int qsim_hip_g02d(qsim_c128 *psik_, double E, double eps, double a1, double a2, size_t nx,
                  size_t ny, qsim_stream_t s)
{
    size_t n = nx * ny;
    NEED(psik_);
    c128 *psik = C(psik_);
    return launch_map(n, s, [=] __device__(size_t idx) {
        double i = (double)fold_half(idx / ny, nx), j = (double)fold_half(idx % ny, ny);
        double t = i * i * a1 + j * j * a2;
        psik[idx] = cdiv(psik[idx], mk(E - t, eps));
    });
}

// op_G03D (src/DeviceTemplate.h:937-976)
int qsim_hip_g03d(qsim_c128 *psik_, double E, double eps, double a1, double a2, double a3,
                  size_t nx, size_t ny, size_t nz, qsim_stream_t s)
{
    size_t n = nx * ny * nz;
    NEED(psik_);
    c128 *psik = C(psik_);
    return launch_map(n, s, [=] __device__(size_t idx) {
        size_t k = idx % nz, r = idx / nz;
        double fk = (double)fold_half(k, nz), fj = (double)fold_half(r % ny, ny),
               fi = (double)fold_half(r / ny, nx);
        double t = fi * fi * a1 + fj * fj * a2 + fk * fk * a3;
        psik[idx] = cdiv(psik[idx], mk(E - t, eps));
    });
}

// op_AddMul: (delpsi + psi0) * v (src/DeviceTemplate.h:1004-1010)
int qsim_hip_add_mul(qsim_c128 *psi_, const qsim_c128 *psi0_, const qsim_c128 *V_, size_t n,
                     qsim_stream_t s)
{
    NEED(psi_);
    NEED(psi0_);
    NEED(V_);
    c128 *psi = C(psi_);
    const c128 *psi0 = C(psi0_), *V = C(V_);
    return launch_map(n, s, [=] __device__(size_t i) { psi[i] = (psi[i] + psi0[i]) * V[i]; });
}

// op_Add3: (ReV + i(ImV+eps)) u + ReV psi0 (src/DeviceTemplate.h:1079-1093)
int qsim_hip_add3(qsim_c128 *out_, const qsim_c128 *dpsi_, const qsim_c128 *psi0_,
                  const qsim_c128 *V_, double eps, size_t n, qsim_stream_t s)
{
    NEED(out_);
    NEED(dpsi_);
    NEED(psi0_);
    NEED(V_);
    c128 *out = C(out_);
    const c128 *dpsi = C(dpsi_), *psi0 = C(psi0_), *V = C(V_);
    return launch_map(n, s, [=] __device__(size_t i) {
        c128 v = V[i];
        out[i] = mk(v.x, v.y + eps) * dpsi[i] + v.x * psi0[i];
    });
}

// op_LinearUpdate: u(1-gamma) + v gamma (src/DeviceTemplate.h:531-542)
int qsim_hip_linear_update(qsim_c128 *psi_, const qsim_c128 *np_, double gamma, size_t n,
                           qsim_stream_t s)
{
    NEED(psi_);
    NEED(np_);
    c128 *psi = C(psi_);
    const c128 *np = C(np_);
    return launch_map(n, s, [=] __device__(size_t i) {
        psi[i] = psi[i] * (1 - gamma) + np[i] * gamma;
    });
}

int qsim_hip_vellekoop(qsim_c128 *psi, const qsim_c128 *np, const qsim_c128 *V, double slow,
                       double eps, size_t n, qsim_stream_t s)
{
    return precond_update<1>(psi, np, V, slow, eps, n, s);
}
int qsim_hip_hao1(qsim_c128 *psi, const qsim_c128 *np, const qsim_c128 *V, double slow,
                  double eps, size_t n, qsim_stream_t s)
{
    return precond_update<2>(psi, np, V, slow, eps, n, s);
}
int qsim_hip_hao2(qsim_c128 *psi, const qsim_c128 *np, const qsim_c128 *V, double slow,
                  double eps, size_t n, qsim_stream_t s)
{
    return precond_update<3>(psi, np, V, slow, eps, n, s);
}

// ------------------------------------------------------------ reductions
#define OUT_CHECK QSIM_CHECK(out != nullptr, "null result pointer")

int qsim_hip_norm2(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t i, double *v) { v[0] = abs2(a[i]); });
}

int qsim_hip_norm2_real(const double *a, size_t n, double *out, qsim_stream_t s)
{
    NEED(a);
    OUT_CHECK;
    return reduce_sum<1>(n, s, out, [=] __device__(size_t i, double *v) { v[0] = a[i] * a[i]; });
}

int qsim_hip_minus_norm2(const qsim_c128 *a_, const qsim_c128 *b_, size_t n, double *out,
                         qsim_stream_t s)
{
    NEED(a_);
    NEED(b_);
    OUT_CHECK;
    const c128 *a = C(a_), *b = C(b_);
    return reduce_sum<1>(n, s, out,
                         [=] __device__(size_t i, double *v) { v[0] = abs2(a[i] - b[i]); });
}

int qsim_hip_max_abs(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    double m2 = 0;
    QSIM_TRY(reduce_ext<true>(n, s, 0.0, &m2, [=] __device__(size_t i) { return abs2(a[i]); }));
    *out = std::sqrt(m2);
    return QSIM_OK;
}

int qsim_hip_min_abs(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    double m2 = 0;
    QSIM_TRY(reduce_ext<false>(n, s, 1.79769313486231570815e+308, &m2,
                               [=] __device__(size_t i) { return abs2(a[i]); }));
    *out = std::sqrt(m2);
    return QSIM_OK;
}

int qsim_hip_sum_real(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t i, double *v) { v[0] = a[i].x; });
}

int qsim_hip_sum_imag(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t i, double *v) { v[0] = a[i].y; });
}

int qsim_hip_abs2_idx(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out,
                         [=] __device__(size_t i, double *v) { v[0] = abs2(a[i]) * (double)i; });
}

int qsim_hip_abs2_idx2(const qsim_c128 *a_, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t i, double *v) {
        double k = (double)i;
        v[0] = abs2(a[i]) * (k * k);
    });
}

int qsim_hip_abs2_idx_real(const double *a, size_t n, double *out, qsim_stream_t s)
{
    NEED(a);
    OUT_CHECK;
    return reduce_sum<1>(n, s, out,
                         [=] __device__(size_t i, double *v) { v[0] = a[i] * a[i] * (double)i; });
}

int qsim_hip_abs2_k1d(const qsim_c128 *a_, double ax, size_t n, double *out, qsim_stream_t s)
{
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t i, double *v) {
        double k = (double)fold_half(i, n);
        v[0] = abs2(a[i]) * (k * k) * ax;
    });
}

int qsim_hip_abs2_k2d(const qsim_c128 *a_, double ax, size_t nx, double ay, size_t ny,
                      double *out, qsim_stream_t s)
{
    size_t n = nx * ny;
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t idx, double *v) {
        double i = (double)fold_half(idx / ny, nx), j = (double)fold_half(idx % ny, ny);
        v[0] = abs2(a[idx]) * (i * i * ax + j * j * ay);
    });
}

int qsim_hip_abs2_k3d(const qsim_c128 *a_, double ax, size_t nx, double ay, size_t ny,
                      double az, size_t nz, double *out, qsim_stream_t s)
{
    size_t n = nx * ny * nz;
    NEED(a_);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out, [=] __device__(size_t idx, double *v) {
        size_t k = idx % nz, r = idx / nz;
        double fk = (double)fold_half(k, nz), fj = (double)fold_half(r % ny, ny),
               fi = (double)fold_half(r / ny, nx);
        v[0] = abs2(a[idx]) * (fi * fi * ax + fj * fj * ay + fk * fk * az);
    });
}

int qsim_hip_abs2_mul_real(const qsim_c128 *a_, const double *V, size_t n, double *out,
                           qsim_stream_t s)
{
    NEED(a_);
    NEED(V);
    OUT_CHECK;
    const c128 *a = C(a_);
    return reduce_sum<1>(n, s, out,
                         [=] __device__(size_t i, double *v) { v[0] = abs2(a[i]) * V[i]; });
}

int qsim_hip_abs2_mul_complex(const qsim_c128 *a_, const qsim_c128 *V_, size_t n,
                              qsim_c128 *out, qsim_stream_t s)
{
    NEED(a_);
    NEED(V_);
    OUT_CHECK;
    const c128 *a = C(a_), *V = C(V_);
    double r[2];
    QSIM_TRY(reduce_sum<2>(n, s, r, [=] __device__(size_t i, double *v) {
        double m = abs2(a[i]);
        v[0] = m * V[i].x;
        v[1] = m * V[i].y;
    }));
    out->re = r[0];
    out->im = r[1];
    return QSIM_OK;
}

// op_Xsec3D (src/DeviceTemplate.h:1104-1134)
int qsim_hip_xsection3d(const qsim_c128 *psi0_, const qsim_c128 *psi_, const qsim_c128 *V_,
                        double kx, double ky, double kz, double x0, double y0, double z0,
                        double dx, double dy, double dz, size_t nx, size_t ny, size_t nz,
                        qsim_c128 *out, qsim_stream_t s)
{
    size_t n = nx * ny * nz;
    NEED(psi0_);
    NEED(psi_);
    NEED(V_);
    OUT_CHECK;
    const c128 *psi0 = C(psi0_), *psi = C(psi_), *V = C(V_);
    double r[2];
    QSIM_TRY(reduce_sum<2>(n, s, r, [=] __device__(size_t idx, double *v) {
        size_t iz = idx % nz, q = idx / nz, iy = q % ny, ix = q / ny;
        double ph = (x0 + (double)ix * dx) * kx + (y0 + (double)iy * dy) * ky +
                    (z0 + (double)iz * dz) * kz;
        double sn, cs;
        sincos(-ph, &sn, &cs);
        c128 t = ((psi0[idx] + psi[idx]) * V[idx].x) * mk(cs, sn);
        v[0] = t.x;
        v[1] = t.y;
    }));
    out->re = r[0];
    out->im = r[1];
    return QSIM_OK;
}

// 2D cross-section sum (src/ScatteringSolverImpl.cpp:186-191), on the device
int qsim_hip_xsection2d(const qsim_c128 *psi0_, const qsim_c128 *psi_, const qsim_c128 *V_,
                        double kx, double ky, double x0, double y0, double dx, double dy,
                        size_t nx, size_t ny, qsim_c128 *out, qsim_stream_t s)
{
    size_t n = nx * ny;
    NEED(psi0_);
    NEED(psi_);
    NEED(V_);
    OUT_CHECK;
    const c128 *psi0 = C(psi0_), *psi = C(psi_), *V = C(V_);
    double r[2];
    QSIM_TRY(reduce_sum<2>(n, s, r, [=] __device__(size_t idx, double *v) {
        size_t ix = idx / ny, iy = idx % ny;
        double x = x0 + dx * (double)ix, y = y0 + dy * (double)iy;
        double ph = kx * x + ky * y;
        double sn, cs;
        sincos(ph, &sn, &cs);
        c128 t = ((psi0[idx] + psi[idx]) * V[idx].x) * mk(cs, -sn);
        v[0] = t.x;
        v[1] = t.y;
    }));
    out->re = r[0];
    out->im = r[1];
    return QSIM_OK;
}

}  // extern "C"
