// hip_device.h -- HipDevice: the reference's Device interface implemented on
// the MI355X through libqsim_hip.so's C ABI (include/qsim_hip.h), plus the HIP
// FourierTransform{1,2,3}D plug-in.  Host C++ only: no HIP headers.
//
// Replaces, at the same plug-in point, DeviceTemplate<CudaExec>
// (src/CudaDevice.cu:33-341) and the cuFFT transforms
// (src/FourierTransform.cpp:84-128, 336-360, 495-519).  Every call is
// synchronous on return where the reference's are (reductions, element
// reads, copies); pointwise ops are enqueued on the device's stream in order.
#pragma once

#include <stdexcept>
#include <string>

#include "device_iface.h"
#include "qsim_hip.h"

namespace qsim_host {

// status -> std::runtime_error, as the reference backends throw
void check(int rc, const char *what);

struct HipDevice : Device {
    explicit HipDevice(int device = 0);

    bool OnMainMem() override { return false; }
    void *AllocMem(size_t bytes) override;
    void Free(void *p) override;
    void MemToDevice(void *out, void const *in, size_t bytes) override;
    void MemToHost(void *out, void const *in, size_t bytes) override;

    void SetZero(DComplex *in, size_t n) override;
    void SetOne(DComplex *in, size_t n) override;
    DReal Abs2Idx(DComplex const *in, size_t n) override;
    DReal Abs2Idx(DReal const *in, size_t n) override;
    void Set(DComplex *in, size_t idx, DComplex alpha) override;
    void Set(DReal *in, size_t idx, DReal alpha) override;
    DComplex At(DComplex const *in, size_t idx, size_t n) override;
    DReal At(DReal const *in, size_t idx, size_t n) override;
    DReal Min(DComplex const *in, size_t n) override;
    DReal Max(DComplex const *in, size_t n) override;
    DReal SumReal(DComplex const *in, size_t n) override;
    DReal SumImag(DComplex const *in, size_t n) override;
    DReal MinusNorm2(DComplex const *in1, DComplex const *in2, size_t n) override;
    DReal Norm2(DReal const *in, size_t nx) override;
    DReal Norm2(DComplex const *in, size_t nx) override;
    void Exp(DComplex *to, DComplex const *from, DComplex alpha, size_t nx) override;
    void Exp(DComplex *to, DReal const *from, DComplex alpha, size_t nx) override;
    void ExpI(DComplex *to, DReal const *from, DReal alpha, size_t nx) override;
    void Copy(DComplex *to, DComplex const *from, size_t n) override;
    void Copy(DReal *to, DReal const *from, size_t n) override;
    void Scale(double *fromto, double scale, size_t nx) override;
    void Scale(DComplex *fromto, double scale, size_t nx) override;
    void Scale(DComplex *fromto, DComplex scale, size_t nx) override;
    void Scale(DReal *to, DReal const *from, DReal scale, size_t nx) override;
    void Scale(DComplex *to, DComplex const *from, double scale, size_t nx) override;
    void Scale(DComplex *to, DComplex const *from, DComplex scale, size_t nx) override;
    void Sub(DComplex *to, DComplex const *in1, DComplex const *in2, size_t n) override;
    void Mul(DComplex *to, DComplex const *from1, DComplex const *from2, size_t nx) override;
    void MulR(DComplex *to, DComplex const *from1, DComplex const *from2, size_t nx) override;
    void MulMinusI(DComplex *to, DComplex const *from1, DComplex const *from2, size_t nx) override;
    void Mul(DComplex *to, DReal const *from1, DComplex const *from2, size_t nx) override;
    void Mul(DComplex *to, DReal const *from1, DComplex alpha, DComplex const *from2,
             size_t nx) override;
    void Mul(DComplex *to, DComplex const *from2, size_t nx) override;
    void MulExp(DComplex *to, DComplex const *from1, DComplex alpha, DComplex const *from2,
                size_t nx) override;
    void MulExp(DComplex *fromto, DComplex alpha, DComplex const *from2, size_t nx) override;
    void MulExp(DComplex *fromto, DComplex alpha, DReal const *from2, size_t nx) override;
    void MulExpI(DComplex *fromto, DReal alpha, DReal const *from2, size_t nx) override;
    void CopyReverseMinu(DComplex *to, DComplex const *from, size_t nx) override;
    DReal Abs2Idx2(DComplex const *psi, size_t nx) override;
    void MulExpIdx2(DComplex *psi, DComplex alpha, size_t nx) override;
    DReal Abs2K1D(DComplex const *psi, DReal ax, size_t nx) override;
    DReal Abs2K2D(DComplex const *psi, DReal ax, size_t nx, DReal ay, size_t ny) override;
    DReal Abs2K3D(DComplex const *psi, DReal ax, size_t nx, DReal ay, size_t ny, DReal az,
                  size_t nz) override;
    void MulExpK1D(DComplex *psi, DComplex ax, size_t nx) override;
    void MulExpK2D(DComplex *psi, DComplex alpha, DReal ax, size_t nx, DReal ay,
                   size_t ny) override;
    void MulExpK3D(DComplex *psi, DComplex alpha, DReal ax, size_t nx, DReal ay, size_t ny,
                   DReal az, size_t nz) override;
    DComplex Abs2Mul(DComplex const *in1, DComplex const *in2, size_t nx) override;
    DReal Abs2Mul(DComplex const *in1, DReal const *in2, size_t nx) override;
    void G01D(DComplex *psik, DReal E, DReal epsilon, DReal alpha, size_t n) override;
    void G02D(DComplex *psik, DReal E, DReal epsilon, DReal a1, DReal a2, size_t nx,
              size_t ny) override;
    void G03D(DComplex *psik, DReal E, DReal epsilon, DReal a1, DReal a2, DReal a3, size_t nx,
              size_t ny, size_t nz) override;
    DComplex Xsection3D(DComplex const *psi0, DComplex const *psi, DComplex const *v, DReal kx,
                        DReal ky, DReal kz, DReal x0, DReal y0, DReal z0, DReal dx, DReal dy,
                        DReal dz, size_t nx, size_t ny, size_t nz) override;
    void AddMul(DComplex *psi, DComplex const *psi0, DComplex const *v, size_t n) override;
    void Add3(DComplex *deltaPsix, DComplex const *psi0x, DComplex const *V, DReal epsilon,
              size_t n) override;
    void Add3(DComplex *add3, DComplex const *deltaPsix, DComplex const *psi0x,
              DComplex const *V, DReal epsilon, size_t n) override;
    void LinearUpdate(DComplex *psi, DComplex const *newPsi, DReal gamma, size_t n) override;
    void Vellekoop(DComplex *psi, DComplex const *newPsi, DComplex const *V, DReal slow,
                   DReal epsilon, size_t n) override;
    void Hao1(DComplex *psi, DComplex const *newPsi, DComplex const *V, DReal slow,
              DReal epsilon, size_t n) override;
    void Hao2(DComplex *psi, DComplex const *newPsi, DComplex const *V, DReal slow,
              DReal epsilon, size_t n) override;

    qsim_stream_t stream = nullptr;
};

// FourierTransform plug-ins (out of place, unnormalised, complex128)
struct HipFourierTransform1D : FourierTransform1D {
    HipFourierTransform1D(size_t n, bool inverse);
    ~HipFourierTransform1D() override;
    void Transform(std::complex<double> const *from, std::complex<double> *to) override;
    qsim_fft_plan *plan = nullptr;
};
struct HipFourierTransform2D : FourierTransform2D {
    HipFourierTransform2D(size_t rows, size_t cols, bool inverse);
    ~HipFourierTransform2D() override;
    void Transform(std::complex<double> const *from, std::complex<double> *to) override;
    qsim_fft_plan *plan = nullptr;
};
struct HipFourierTransform3D : FourierTransform3D {
    HipFourierTransform3D(size_t n1, size_t n2, size_t n3, bool inverse);
    ~HipFourierTransform3D() override;
    void Transform(std::complex<double> const *from, std::complex<double> *to) override;
    qsim_fft_plan *plan = nullptr;
};

// Device::Create(DeviceType::GPU_HIP) hook for the reference factory
// (src/Device.cpp:6-38).
Device *CreateHipDevice();

}  // namespace qsim_host
