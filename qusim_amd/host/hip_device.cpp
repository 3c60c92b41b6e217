// hip_device.cpp -- HipDevice / HipFourierTransform*D over the C ABI.
#include "hip_device.h"

#include <cstring>

namespace qsim_host {

void check(int rc, const char *what)
{
    if (rc != QSIM_OK)
        throw std::runtime_error(std::string(what) + ": " + qsim_hip_last_error());
}

namespace {
inline qsim_c128 *C(Device::DComplex *p) { return reinterpret_cast<qsim_c128 *>(p); }
inline const qsim_c128 *C(Device::DComplex const *p)
{
    return reinterpret_cast<const qsim_c128 *>(p);
}
inline qsim_c128 V(Device::DComplex z) { return qsim_c128{z.real(), z.imag()}; }
}  // namespace

#define Q(call) check((call), #call)

HipDevice::HipDevice(int device) { Q(qsim_hip_init(device)); }

void *HipDevice::AllocMem(size_t bytes)
{
    void *p = nullptr;
    int rc = qsim_hip_malloc(&p, bytes);
    if (rc != QSIM_OK)  // as the reference allocator (src/DeviceTemplate.h:327-334)
        throw std::runtime_error("can't alloc " + std::to_string(bytes) + " bytes");
    return p;
}
void HipDevice::Free(void *p) { Q(qsim_hip_free(p)); }
void HipDevice::MemToDevice(void *out, void const *in, size_t bytes)
{
    Q(qsim_hip_synchronize(stream));
    Q(qsim_hip_memcpy_h2d(out, in, bytes));
}
void HipDevice::MemToHost(void *out, void const *in, size_t bytes)
{
    Q(qsim_hip_synchronize(stream));
    Q(qsim_hip_memcpy_d2h(out, in, bytes));
}

void HipDevice::SetZero(DComplex *in, size_t n) { Q(qsim_hip_set_zero(C(in), n, stream)); }
void HipDevice::SetOne(DComplex *in, size_t n) { Q(qsim_hip_set_one(C(in), n, stream)); }

HipDevice::DReal HipDevice::Abs2Idx(DComplex const *in, size_t n)
{
    double r;
    Q(qsim_hip_abs2_idx(C(in), n, &r, stream));
    return r;
}
HipDevice::DReal HipDevice::Abs2Idx(DReal const *in, size_t n)
{
    double r;
    Q(qsim_hip_abs2_idx_real(in, n, &r, stream));
    return r;
}
void HipDevice::Set(DComplex *in, size_t idx, DComplex a) { Q(qsim_hip_set_elem(C(in), idx, V(a), stream)); }
void HipDevice::Set(DReal *in, size_t idx, DReal a) { Q(qsim_hip_set_elem_real(in, idx, a, stream)); }
HipDevice::DComplex HipDevice::At(DComplex const *in, size_t idx, size_t)
{
    qsim_c128 r;
    Q(qsim_hip_get_elem(C(in), idx, &r, stream));
    return DComplex(r.re, r.im);
}
HipDevice::DReal HipDevice::At(DReal const *in, size_t idx, size_t)
{
    double r;
    Q(qsim_hip_get_elem_real(in, idx, &r, stream));
    return r;
}

#define RED1(NAME, ...)                    \
    double r_;                             \
    Q(NAME(__VA_ARGS__, &r_, stream));     \
    return r_;

HipDevice::DReal HipDevice::Min(DComplex const *in, size_t n) { RED1(qsim_hip_min_abs, C(in), n) }
HipDevice::DReal HipDevice::Max(DComplex const *in, size_t n) { RED1(qsim_hip_max_abs, C(in), n) }
HipDevice::DReal HipDevice::SumReal(DComplex const *in, size_t n) { RED1(qsim_hip_sum_real, C(in), n) }
HipDevice::DReal HipDevice::SumImag(DComplex const *in, size_t n) { RED1(qsim_hip_sum_imag, C(in), n) }
HipDevice::DReal HipDevice::MinusNorm2(DComplex const *a, DComplex const *b, size_t n)
{
    RED1(qsim_hip_minus_norm2, C(a), C(b), n)
}
HipDevice::DReal HipDevice::Norm2(DReal const *in, size_t n) { RED1(qsim_hip_norm2_real, in, n) }
HipDevice::DReal HipDevice::Norm2(DComplex const *in, size_t n) { RED1(qsim_hip_norm2, C(in), n) }

void HipDevice::Exp(DComplex *to, DComplex const *from, DComplex a, size_t n)
{
    Q(qsim_hip_exp_complex(C(to), C(from), V(a), n, stream));
}
void HipDevice::Exp(DComplex *to, DReal const *from, DComplex a, size_t n)
{
    Q(qsim_hip_exp_real(C(to), from, V(a), n, stream));
}
void HipDevice::ExpI(DComplex *to, DReal const *from, DReal a, size_t n)
{
    Q(qsim_hip_exp_i_real(C(to), from, a, n, stream));
}
void HipDevice::Copy(DComplex *to, DComplex const *from, size_t n) { Q(qsim_hip_copy(C(to), C(from), n, stream)); }
void HipDevice::Copy(DReal *to, DReal const *from, size_t n) { Q(qsim_hip_copy_real(to, from, n, stream)); }
void HipDevice::Scale(double *a, double s, size_t n) { Q(qsim_hip_scale_real(a, s, n, stream)); }
void HipDevice::Scale(DComplex *a, double s, size_t n) { Q(qsim_hip_scale(C(a), s, n, stream)); }
void HipDevice::Scale(DComplex *a, DComplex s, size_t n) { Q(qsim_hip_scale_c(C(a), V(s), n, stream)); }
void HipDevice::Scale(DReal *to, DReal const *from, DReal s, size_t n)
{
    Q(qsim_hip_scale_real_to(to, from, s, n, stream));
}
void HipDevice::Scale(DComplex *to, DComplex const *from, double s, size_t n)
{
    Q(qsim_hip_scale_to(C(to), C(from), s, n, stream));
}
void HipDevice::Scale(DComplex *to, DComplex const *from, DComplex s, size_t n)
{
    Q(qsim_hip_scale_c_to(C(to), C(from), V(s), n, stream));
}
void HipDevice::Sub(DComplex *to, DComplex const *a, DComplex const *b, size_t n)
{
    Q(qsim_hip_sub(C(to), C(a), C(b), n, stream));
}
void HipDevice::Mul(DComplex *to, DComplex const *a, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul3(C(to), C(a), C(b), n, stream));
}
void HipDevice::MulR(DComplex *to, DComplex const *a, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul_r(C(to), C(a), C(b), n, stream));
}
void HipDevice::MulMinusI(DComplex *to, DComplex const *a, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul_minus_i(C(to), C(a), C(b), n, stream));
}
void HipDevice::Mul(DComplex *to, DReal const *a, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul_real_c(C(to), a, C(b), n, stream));
}
void HipDevice::Mul(DComplex *to, DReal const *a, DComplex alpha, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul_real_alpha_c(C(to), a, V(alpha), C(b), n, stream));
}
void HipDevice::Mul(DComplex *to, DComplex const *from2, size_t n)
{
    Q(qsim_hip_mul(C(to), C(from2), n, stream));
}
void HipDevice::MulExp(DComplex *to, DComplex const *a, DComplex alpha, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul_exp3(C(to), C(a), V(alpha), C(b), n, stream));
}
void HipDevice::MulExp(DComplex *fromto, DComplex alpha, DComplex const *b, size_t n)
{
    Q(qsim_hip_mul_exp_complex(C(fromto), V(alpha), C(b), n, stream));
}
void HipDevice::MulExp(DComplex *fromto, DComplex alpha, DReal const *b, size_t n)
{
    Q(qsim_hip_mul_exp_real(C(fromto), V(alpha), b, n, stream));
}
void HipDevice::MulExpI(DComplex *fromto, DReal alpha, DReal const *b, size_t n)
{
    Q(qsim_hip_mul_exp_i_real(C(fromto), alpha, b, n, stream));
}
void HipDevice::CopyReverseMinu(DComplex *to, DComplex const *from, size_t n)
{
    Q(qsim_hip_copy_reverse_minus(C(to), C(from), n, stream));
}
HipDevice::DReal HipDevice::Abs2Idx2(DComplex const *psi, size_t n) { RED1(qsim_hip_abs2_idx2, C(psi), n) }
void HipDevice::MulExpIdx2(DComplex *psi, DComplex alpha, size_t n)
{
    Q(qsim_hip_mul_exp_idx2(C(psi), V(alpha), n, stream));
}
HipDevice::DReal HipDevice::Abs2K1D(DComplex const *psi, DReal ax, size_t nx)
{
    RED1(qsim_hip_abs2_k1d, C(psi), ax, nx)
}
HipDevice::DReal HipDevice::Abs2K2D(DComplex const *psi, DReal ax, size_t nx, DReal ay, size_t ny)
{
    RED1(qsim_hip_abs2_k2d, C(psi), ax, nx, ay, ny)
}
HipDevice::DReal HipDevice::Abs2K3D(DComplex const *psi, DReal ax, size_t nx, DReal ay, size_t ny,
                                    DReal az, size_t nz)
{
    RED1(qsim_hip_abs2_k3d, C(psi), ax, nx, ay, ny, az, nz)
}
void HipDevice::MulExpK1D(DComplex *psi, DComplex ax, size_t nx)
{
    Q(qsim_hip_mul_exp_k1d(C(psi), V(ax), nx, stream));
}
void HipDevice::MulExpK2D(DComplex *psi, DComplex alpha, DReal ax, size_t nx, DReal ay, size_t ny)
{
    Q(qsim_hip_mul_exp_k2d(C(psi), V(alpha), ax, nx, ay, ny, stream));
}
void HipDevice::MulExpK3D(DComplex *psi, DComplex alpha, DReal ax, size_t nx, DReal ay, size_t ny,
                          DReal az, size_t nz)
{
    Q(qsim_hip_mul_exp_k3d(C(psi), V(alpha), ax, nx, ay, ny, az, nz, stream));
}
HipDevice::DComplex HipDevice::Abs2Mul(DComplex const *a, DComplex const *b, size_t n)
{
    qsim_c128 r;
    Q(qsim_hip_abs2_mul_complex(C(a), C(b), n, &r, stream));
    return DComplex(r.re, r.im);
}
HipDevice::DReal HipDevice::Abs2Mul(DComplex const *a, DReal const *b, size_t n)
{
    RED1(qsim_hip_abs2_mul_real, C(a), b, n)
}
void HipDevice::G01D(DComplex *psik, DReal E, DReal eps, DReal a, size_t n)
{
    Q(qsim_hip_g01d(C(psik), E, eps, a, n, stream));
}
void HipDevice::G02D(DComplex *psik, DReal E, DReal eps, DReal a1, DReal a2, size_t nx, size_t ny)
{
    Q(qsim_hip_g02d(C(psik), E, eps, a1, a2, nx, ny, stream));
}
void HipDevice::G03D(DComplex *psik, DReal E, DReal eps, DReal a1, DReal a2, DReal a3, size_t nx,
                     size_t ny, size_t nz)
{
    Q(qsim_hip_g03d(C(psik), E, eps, a1, a2, a3, nx, ny, nz, stream));
}
HipDevice::DComplex HipDevice::Xsection3D(DComplex const *psi0, DComplex const *psi,
                                          DComplex const *v, DReal kx, DReal ky, DReal kz,
                                          DReal x0, DReal y0, DReal z0, DReal dx, DReal dy,
                                          DReal dz, size_t nx, size_t ny, size_t nz)
{
    qsim_c128 r;
    Q(qsim_hip_xsection3d(C(psi0), C(psi), C(v), kx, ky, kz, x0, y0, z0, dx, dy, dz, nx, ny, nz,
                          &r, stream));
    return DComplex(r.re, r.im);
}
void HipDevice::AddMul(DComplex *psi, DComplex const *psi0, DComplex const *v, size_t n)
{
    Q(qsim_hip_add_mul(C(psi), C(psi0), C(v), n, stream));
}
void HipDevice::Add3(DComplex *dpsi, DComplex const *psi0, DComplex const *v, DReal eps, size_t n)
{
    Q(qsim_hip_add3(C(dpsi), C(dpsi), C(psi0), C(v), eps, n, stream));
}
void HipDevice::Add3(DComplex *add3, DComplex const *dpsi, DComplex const *psi0,
                     DComplex const *v, DReal eps, size_t n)
{
    Q(qsim_hip_add3(C(add3), C(dpsi), C(psi0), C(v), eps, n, stream));
}
void HipDevice::LinearUpdate(DComplex *psi, DComplex const *np, DReal gamma, size_t n)
{
    Q(qsim_hip_linear_update(C(psi), C(np), gamma, n, stream));
}
void HipDevice::Vellekoop(DComplex *psi, DComplex const *np, DComplex const *v, DReal slow,
                          DReal eps, size_t n)
{
    Q(qsim_hip_vellekoop(C(psi), C(np), C(v), slow, eps, n, stream));
}
void HipDevice::Hao1(DComplex *psi, DComplex const *np, DComplex const *v, DReal slow, DReal eps,
                     size_t n)
{
    Q(qsim_hip_hao1(C(psi), C(np), C(v), slow, eps, n, stream));
}
void HipDevice::Hao2(DComplex *psi, DComplex const *np, DComplex const *v, DReal slow, DReal eps,
                     size_t n)
{
    Q(qsim_hip_hao2(C(psi), C(np), C(v), slow, eps, n, stream));
}

// ------------------------------------------------------------------ FFT
HipFourierTransform1D::HipFourierTransform1D(size_t n, bool inverse)
{
    Q(qsim_hip_fft_plan_create(&plan, 1, &n, inverse ? 1 : 0, 64));
}
HipFourierTransform1D::~HipFourierTransform1D() { qsim_hip_fft_plan_destroy(plan); }
void HipFourierTransform1D::Transform(std::complex<double> const *from, std::complex<double> *to)
{
    Q(qsim_hip_fft_exec(plan, from, to, nullptr));
}

HipFourierTransform2D::HipFourierTransform2D(size_t rows, size_t cols, bool inverse)
{
    size_t d[2] = {rows, cols};
    Q(qsim_hip_fft_plan_create(&plan, 2, d, inverse ? 1 : 0, 64));
}
HipFourierTransform2D::~HipFourierTransform2D() { qsim_hip_fft_plan_destroy(plan); }
void HipFourierTransform2D::Transform(std::complex<double> const *from, std::complex<double> *to)
{
    Q(qsim_hip_fft_exec(plan, from, to, nullptr));
}

HipFourierTransform3D::HipFourierTransform3D(size_t n1, size_t n2, size_t n3, bool inverse)
{
    size_t d[3] = {n1, n2, n3};
    Q(qsim_hip_fft_plan_create(&plan, 3, d, inverse ? 1 : 0, 64));
}
HipFourierTransform3D::~HipFourierTransform3D() { qsim_hip_fft_plan_destroy(plan); }
void HipFourierTransform3D::Transform(std::complex<double> const *from, std::complex<double> *to)
{
    Q(qsim_hip_fft_exec(plan, from, to, nullptr));
}

Device *CreateHipDevice() { return new HipDevice(0); }

}  // namespace qsim_host
