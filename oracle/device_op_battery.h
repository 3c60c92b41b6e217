// device_op_battery.h -- TEST INFRASTRUCTURE: every elementwise op and
// reduction of the Device interface (src/Device.h:82-173) run once on fixed
// inputs, through the interface's own allocation and copy methods.  The same
// template is compiled against the reference's CpuSeq device
// (src/DeviceTemplate.h, oracle/ref_driver.cpp: golden vectors) and against
// HipDevice (tests/cpp/hip_device_driver.cpp: the GPU side), so each op of the
// HIP device is checked against the reference's own functor.
//
// Inputs: complex a, b, c and real r of n = nx*ny*nz points (r > 0).
// Output: kBatteryArrays arrays of n complex values (real results as
// (x, 0)) followed by kBatteryScalars complex scalars.
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

constexpr int kBatteryArrays = 41;
constexpr int kBatteryScalars = 18;

template <class Dev>
void device_op_battery(Dev &dev, size_t nx, size_t ny, size_t nz, const std::complex<double> *a,
                       const std::complex<double> *b, const std::complex<double> *c,
                       const double *r, std::complex<double> *out)
{
    using C = std::complex<double>;
    const size_t n = nx * ny * nz;
    auto allocc = [&]() { return (C *)dev.AllocMem(n * sizeof(C)); };
    auto allocr = [&]() { return (double *)dev.AllocMem(n * sizeof(double)); };
    C *da = allocc(), *db = allocc(), *dc = allocc(), *w = allocc();
    double *dr = allocr(), *wr = allocr();
    dev.MemToDevice(da, a, n * sizeof(C));
    dev.MemToDevice(db, b, n * sizeof(C));
    dev.MemToDevice(dc, c, n * sizeof(C));
    dev.MemToDevice(dr, r, n * sizeof(double));
    int slot = 0;
    auto emit = [&](const C *p) { dev.MemToHost(out + n * slot++, p, n * sizeof(C)); };
    auto emitr = [&](const double *p) {
        std::vector<double> h(n);
        dev.MemToHost(h.data(), p, n * sizeof(double));
        for (size_t i = 0; i < n; ++i) out[n * slot + i] = C(h[i], 0);
        ++slot;
    };
    auto fresh = [&]() {
        dev.Copy(w, da, n);
        return w;
    };
    const C al(0.3, -0.7);
    // 0-1
    dev.SetZero(w, n);
    emit(w);
    dev.SetOne(w, n);
    emit(w);
    // 2-4: Exp of complex and real arrays, ExpI
    dev.Exp(w, da, al, n);
    emit(w);
    dev.Exp(w, dr, al, n);
    emit(w);
    dev.ExpI(w, dr, 0.9, n);
    emit(w);
    // 5-9: Copy and the complex Scale forms
    dev.Copy(w, da, n);
    emit(w);
    dev.Scale(fresh(), 1.7, n);
    emit(w);
    dev.Scale(fresh(), C(0.4, -1.1), n);
    emit(w);
    dev.Scale(w, da, 0.6, n);
    emit(w);
    dev.Scale(w, da, C(0.2, 0.9), n);
    emit(w);
    // 10-16: Sub and the Mul forms
    dev.Sub(w, da, db, n);
    emit(w);
    dev.Mul(w, da, db, n);
    emit(w);
    dev.MulR(w, da, db, n);
    emit(w);
    dev.MulMinusI(w, da, db, n);
    emit(w);
    dev.Mul(w, dr, db, n);
    emit(w);
    dev.Mul(w, dr, C(1.3, -0.4), db, n);
    emit(w);
    dev.Mul(fresh(), db, n);
    emit(w);
    // 17-20: the MulExp forms
    const C am(0.1, 0.25);
    dev.MulExp(w, da, am, db, n);
    emit(w);
    dev.MulExp(fresh(), am, db, n);
    emit(w);
    dev.MulExp(fresh(), am, dr, n);
    emit(w);
    dev.MulExpI(fresh(), 0.35, dr, n);
    emit(w);
    // 21: CopyReverseMinu writes backwards from its first argument
    dev.SetZero(w, n);
    dev.CopyReverseMinu(w + n - 1, da, n);
    emit(w);
    // 22-28: index-phase and kinetic factors, Green's functions
    dev.MulExpIdx2(fresh(), C(-0.001, 0.02), n);
    emit(w);
    dev.MulExpK1D(fresh(), C(-0.002, 0.03), n);
    emit(w);
    dev.MulExpK2D(fresh(), C(0.2, 0.1), 0.03, nx * ny, 0.05, nz);
    emit(w);
    dev.MulExpK3D(fresh(), C(0.2, 0.1), 0.03, nx, 0.05, ny, 0.07, nz);
    emit(w);
    dev.G01D(fresh(), 1.3, 0.2, 0.07, n);
    emit(w);
    dev.G02D(fresh(), 1.3, 0.2, 0.07, 0.11, nx * ny, nz);
    emit(w);
    dev.G03D(fresh(), 1.3, 0.2, 0.07, 0.11, 0.05, nx, ny, nz);
    emit(w);
    // 29-35: the Born updates
    dev.AddMul(fresh(), db, dc, n);
    emit(w);
    dev.Add3(fresh(), db, dc, 0.3, n);
    emit(w);
    dev.Add3(w, da, db, dc, 0.3, n);
    emit(w);
    dev.LinearUpdate(fresh(), db, 0.4, n);
    emit(w);
    dev.Vellekoop(fresh(), db, dc, 0.6, 0.5, n);
    emit(w);
    dev.Hao1(fresh(), db, dc, 0.6, 0.5, n);
    emit(w);
    dev.Hao2(fresh(), db, dc, 0.6, 0.5, n);
    emit(w);
    // 36-38: real Copy and Scale
    dev.Copy(wr, dr, n);
    emitr(wr);
    dev.Copy(wr, dr, n);
    dev.Scale(wr, 1.3, n);
    emitr(wr);
    dev.Scale(wr, dr, 0.7, n);
    emitr(wr);
    // 39-40: Set
    dev.Set(fresh(), 3, C(5.0, -6.0));
    emit(w);
    dev.Copy(wr, dr, n);
    dev.Set(wr, 2, 4.5);
    emitr(wr);
    // scalars
    C *s = out + n * kBatteryArrays;
    int k = 0;
    s[k++] = dev.Abs2Idx(da, n);
    s[k++] = dev.Abs2Idx(dr, n);
    s[k++] = dev.At(da, 4, n);
    s[k++] = dev.At(dr, 5, n);
    s[k++] = dev.Min(da, n);
    s[k++] = dev.Max(da, n);
    s[k++] = dev.SumReal(da, n);
    s[k++] = dev.SumImag(da, n);
    s[k++] = dev.MinusNorm2(da, db, n);
    s[k++] = dev.Norm2(dr, n);
    s[k++] = dev.Norm2(da, n);
    s[k++] = dev.Abs2Idx2(da, n);
    s[k++] = dev.Abs2K1D(da, 0.3, n);
    s[k++] = dev.Abs2K2D(da, 0.2, nx * ny, 0.5, nz);
    s[k++] = dev.Abs2K3D(da, 0.2, nx, 0.5, ny, 0.3, nz);
    s[k++] = dev.Abs2Mul(da, db, n);
    s[k++] = dev.Abs2Mul(da, dr, n);
    s[k++] = dev.Xsection3D(da, db, dc, 0.3, -0.2, 0.9, -1.0, -1.5, -2.0, 0.25, 0.3, 0.35, nx, ny,
                            nz);
    for (void *p : {(void *)da, (void *)db, (void *)dc, (void *)w, (void *)dr, (void *)wr})
        dev.Free(p);
}
