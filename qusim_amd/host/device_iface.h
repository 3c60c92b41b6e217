// device_iface.h -- the plug-in interfaces the HIP device implements.
//
// The reference's Device (src/Device.h:37-174) and FourierTransform{1,2,3}D
// (src/FourierTransform.h:12-36) interfaces, restated with the same method
// names, signatures and meaning so HipDevice / HipFourierTransform*D compile
// unchanged inside the reference tree (where src/Device.h and
// src/FourierTransform.h provide them) and standalone here.  The include
// guards match the reference headers' so only one definition is ever seen.
#ifndef QUSIM_DEVICE
#define QUSIM_DEVICE

#include <complex>
#include <cstddef>
#include <memory>

enum class DeviceType {
    CPU_SEQ,
    CPU_PAR,
    CPU_VEC,
    CPU_PAR_VEC,
    GPU_CUDA,
    GPU_HIP,
};

struct Device {
    using DReal = double;
    using DComplex = std::complex<DReal>;

    virtual ~Device() {}

    virtual bool OnMainMem() = 0;
    virtual void *AllocMem(size_t bytes) = 0;
    virtual void Free(void *) = 0;

    template <class T>
    T *Alloc(size_t n) { return (T *)AllocMem(sizeof(T) * n); }
    template <class T>
    void SafeFree(T *&p)
    {
        if (p) {
            Free(p);
            p = nullptr;
        }
    }

    void ToDevice(DComplex *out, DComplex const *in, size_t n) { MemToDevice(out, in, n * sizeof(DComplex)); }
    void ToDevice(DReal *out, DReal const *in, size_t n) { MemToDevice(out, in, n * sizeof(DReal)); }
    void ToHost(DComplex *out, DComplex const *in, size_t n) { MemToHost(out, in, n * sizeof(DComplex)); }
    void ToHost(DReal *out, DReal const *in, size_t n) { MemToHost(out, in, n * sizeof(DReal)); }
    virtual void MemToDevice(void *out, void const *in, size_t bytes) = 0;
    virtual void MemToHost(void *out, void const *in, size_t bytes) = 0;

    virtual void SetZero(DComplex *in, size_t n) = 0;
    virtual void SetOne(DComplex *in, size_t n) = 0;
    virtual DReal Abs2Idx(DComplex const *in, size_t n) = 0;
    virtual DReal Abs2Idx(DReal const *in, size_t n) = 0;
    virtual void Set(DComplex *in, size_t idx, DComplex alpha) = 0;
    virtual void Set(DReal *in, size_t idx, DReal alpha) = 0;
    virtual DComplex At(DComplex const *in, size_t idx, size_t n) = 0;
    virtual DReal At(DReal const *in, size_t idx, size_t n) = 0;
    virtual DReal Min(DComplex const *in, size_t n) = 0;
    virtual DReal Max(DComplex const *in, size_t n) = 0;
    virtual DReal SumReal(DComplex const *in, size_t n) = 0;
    virtual DReal SumImag(DComplex const *in, size_t n) = 0;
    virtual DReal MinusNorm2(DComplex const *in1, DComplex const *in2, size_t n) = 0;
    virtual DReal Norm2(DReal const *in, size_t nx) = 0;
    virtual DReal Norm2(DComplex const *in, size_t nx) = 0;
    virtual void Exp(DComplex *to, DComplex const *from, DComplex alpha, size_t nx) = 0;
    virtual void Exp(DComplex *to, DReal const *from, DComplex alpha, size_t nx) = 0;
    virtual void ExpI(DComplex *to, DReal const *from, DReal alpha, size_t nx) = 0;
    virtual void Copy(DComplex *to, DComplex const *from, size_t n) = 0;
    virtual void Copy(DReal *to, DReal const *from, size_t n) = 0;
    virtual void Scale(double *fromto, double scale, size_t nx) = 0;
    virtual void Scale(DComplex *fromto, double scale, size_t nx) = 0;
    virtual void Scale(DComplex *fromto, DComplex scale, size_t nx) = 0;
    virtual void Scale(DReal *to, DReal const *from, DReal scale, size_t nx) = 0;
    virtual void Scale(DComplex *to, DComplex const *from, double scale, size_t nx) = 0;
    virtual void Scale(DComplex *to, DComplex const *from, DComplex scale, size_t nx) = 0;
    virtual void Sub(DComplex *to, DComplex const *in1, DComplex const *in2, size_t n) = 0;
    virtual void Mul(DComplex *to, DComplex const *from1, DComplex const *from2, size_t nx) = 0;
    virtual void MulR(DComplex *to, DComplex const *from1, DComplex const *from2, size_t nx) = 0;
    virtual void MulMinusI(DComplex *to, DComplex const *from1, DComplex const *from2, size_t nx) = 0;
    virtual void Mul(DComplex *to, DReal const *from1, DComplex const *from2, size_t nx) = 0;
    virtual void Mul(DComplex *to, DReal const *from1, DComplex alpha, DComplex const *from2, size_t nx) = 0;
    virtual void Mul(DComplex *to, DComplex const *from2, size_t nx) = 0;
    virtual void MulExp(DComplex *to, DComplex const *from1, DComplex alpha, DComplex const *from2, size_t nx) = 0;
    virtual void MulExp(DComplex *fromto, DComplex alpha, DComplex const *from2, size_t nx) = 0;
    virtual void MulExp(DComplex *fromto, DComplex alpha, DReal const *from2, size_t nx) = 0;
    virtual void MulExpI(DComplex *fromto, DReal alpha, DReal const *from2, size_t nx) = 0;
    virtual void CopyReverseMinu(DComplex *to, DComplex const *from, size_t nx) = 0;
    virtual DReal Abs2Idx2(DComplex const *psi, size_t nx) = 0;
    virtual void MulExpIdx2(DComplex *psi, DComplex alpha, size_t nx) = 0;
    virtual DReal Abs2K1D(DComplex const *psi, DReal ax, size_t nx) = 0;
    virtual DReal Abs2K2D(DComplex const *psi, DReal ax, size_t nx, DReal ay, size_t ny) = 0;
    virtual DReal Abs2K3D(DComplex const *psi, DReal ax, size_t nx, DReal ay, size_t ny, DReal az, size_t nz) = 0;
    virtual void MulExpK1D(DComplex *psi, DComplex ax, size_t nx) = 0;
    virtual void MulExpK2D(DComplex *psi, DComplex alpha, DReal ax, size_t nx, DReal ay, size_t ny) = 0;
    virtual void MulExpK3D(DComplex *psi, DComplex alpha, DReal ax, size_t nx, DReal ay, size_t ny, DReal az, size_t nz) = 0;
    virtual DComplex Abs2Mul(DComplex const *in1, DComplex const *in2, size_t nx) = 0;
    virtual DReal Abs2Mul(DComplex const *in1, DReal const *in2, size_t nx) = 0;
    virtual void G01D(DComplex *psik, DReal E, DReal epsilon, DReal alpha, size_t n) = 0;
    virtual void G02D(DComplex *psik, DReal E, DReal epsilon, DReal a1, DReal a2, size_t nx, size_t ny) = 0;
    virtual void G03D(DComplex *psik, DReal E, DReal epsilon, DReal a1, DReal a2, DReal a3, size_t nx, size_t ny, size_t nz) = 0;
    virtual DComplex Xsection3D(DComplex const *psi0, DComplex const *psi, DComplex const *v,
                                DReal kx, DReal ky, DReal kz, DReal x0, DReal y0, DReal z0,
                                DReal dx, DReal dy, DReal dz, size_t nx, size_t ny, size_t nz) = 0;
    virtual void AddMul(DComplex *psi, DComplex const *psi0, DComplex const *v, size_t n) = 0;
    virtual void Add3(DComplex *deltaPsix, DComplex const *psi0x, DComplex const *V, DReal epsilon, size_t n) = 0;
    virtual void Add3(DComplex *add3, DComplex const *deltaPsix, DComplex const *psi0x, DComplex const *V, DReal epsilon, size_t n) = 0;
    virtual void LinearUpdate(DComplex *psi, DComplex const *newPsi, DReal gamma, size_t n) = 0;
    virtual void Vellekoop(DComplex *psi, DComplex const *newPsi, DComplex const *V, DReal slow, DReal epsilon, size_t n) = 0;
    virtual void Hao1(DComplex *psi, DComplex const *newPsi, DComplex const *V, DReal slow, DReal epsilon, size_t n) = 0;
    virtual void Hao2(DComplex *psi, DComplex const *newPsi, DComplex const *V, DReal slow, DReal epsilon, size_t n) = 0;
};

#endif  // QUSIM_DEVICE

#ifndef QSIM_FFT_IFACE
#define QSIM_FFT_IFACE
// FourierTransform plug-in (src/FourierTransform.h:12-36).  Inside the
// reference tree the build defines QSIM_FFT_IFACE and includes the reference
// header instead.
struct FourierTransform1D {
    virtual void Transform(std::complex<double> const *from, std::complex<double> *to) = 0;
    virtual ~FourierTransform1D() {}
};
struct FourierTransform2D {
    virtual void Transform(std::complex<double> const *from, std::complex<double> *to) = 0;
    virtual ~FourierTransform2D() {}
};
struct FourierTransform3D {
    virtual void Transform(std::complex<double> const *from, std::complex<double> *to) = 0;
    virtual ~FourierTransform3D() {}
};
#endif
