// ref_driver.cpp -- builds oracle/_ref/libqsim_ref.so from the reference's own
// sources, where they lie under the reference tree (see oracle/Makefile).
//
// TEST INFRASTRUCTURE ONLY.  Compiled from the reference: kissfft.hh (the
// reference CPU FFT), DeviceTemplate.h + Device.cpp (the CpuSeq device ops),
// Perburbation.h (BornSeries / PreconditionalBornSeries / absorbing layers),
// SplittingUtils.{h,cpp} (QuUpdatePsi and the O4 constants).
//
// Not buildable here and therefore restated in this driver, in the op order
// of the cited reference files:
//   * SplittingMethod1D/2D::ExpV/ExpT (src/SplittingMethod1D.cpp:40-89,
//     src/SplittingMethod2D.cpp:38-121): SplittingMethod2D.h includes Eigen,
//     which the image lacks.
//   * KissFourierTransform2D/3D loops (src/FourierTransform.cpp:176-291,
//     :379-453): FourierTransform.cpp links FFTW, which the image lacks.
//   * The Compute() loop (src/Perturbation2DImpl.cpp:52-83,
//     src/Perturbation3DImpl.cpp:71-126): ScatteringSolverImpl.h includes Eigen.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "kissfft.hh"
#include "DeviceTemplate.h"
#include "FourierTransform.h"
#include "Perburbation.h"
#include "SplittingUtils.h"
#include "Utils.h"

#include "device_op_battery.h"

using Dev = DeviceTemplate<std::execution::sequenced_policy>;
using Cx = std::complex<double>;

namespace {

// KissFourierTransform2D::Transform (src/FourierTransform.cpp:248-284): rows
// (inner axis, length ny) first, then columns.
struct Kiss2D : FourierTransform2D {
    size_t nx, ny;
    kissfft<double> fx, fy;
    Kiss2D(size_t nx_, size_t ny_, bool inv) : nx(nx_), ny(ny_), fx(nx_, inv), fy(ny_, inv) {}
    void Transform(Cx const* from, Cx* to) override
    {
        for (size_t i = 0; i < nx; ++i) fy.transform(from + i * ny, to + i * ny);
        std::vector<Cx> a(nx), b(nx);
        for (size_t j = 0; j < ny; ++j) {
            for (size_t i = 0; i < nx; ++i) a[i] = to[i * ny + j];
            fx.transform(a.data(), b.data());
            for (size_t i = 0; i < nx; ++i) to[i * ny + j] = b[i];
        }
    }
};

// KissFourierTransform3D::Transform (src/FourierTransform.cpp:404-447): z, y, x.
struct Kiss3D : FourierTransform3D {
    size_t nx, ny, nz;
    kissfft<double> fx, fy, fz;
    Kiss3D(size_t nx_, size_t ny_, size_t nz_, bool inv)
        : nx(nx_), ny(ny_), nz(nz_), fx(nx_, inv), fy(ny_, inv), fz(nz_, inv) {}
    void Transform(Cx const* from, Cx* to) override
    {
        for (size_t i = 0; i < nx; ++i)
            for (size_t j = 0; j < ny; ++j) {
                size_t idx = (i * ny + j) * nz;
                fz.transform(from + idx, to + idx);
            }
        std::vector<Cx> a(ny), b(ny);
        for (size_t i = 0; i < nx; ++i)
            for (size_t k = 0; k < nz; ++k) {
                for (size_t j = 0; j < ny; ++j) a[j] = to[(i * ny + j) * nz + k];
                fy.transform(a.data(), b.data());
                for (size_t j = 0; j < ny; ++j) to[(i * ny + j) * nz + k] = b[j];
            }
        std::vector<Cx> c(nx), d(nx);
        for (size_t j = 0; j < ny; ++j)
            for (size_t k = 0; k < nz; ++k) {
                for (size_t i = 0; i < nx; ++i) c[i] = to[(i * ny + j) * nz + k];
                fx.transform(c.data(), d.data());
                for (size_t i = 0; i < nx; ++i) to[(i * ny + j) * nz + k] = d[i];
            }
    }
};

// Split-operator 1D, driven through QuUpdatePsi (src/SplittingUtils.h:14-64).
// fVc (complex potential) replaces fV when set: ExpV then runs the
// reference's complex MulExp (src/Device.h:129, src/DeviceTemplate.h:1478-1486)
// instead of the real one the evolver uses (src/SplittingMethod1D.cpp:84-89).
struct Split1D {
    Dev dev;
    Cx* fPsi = nullptr;
    double* fV = nullptr;
    Cx* fVc = nullptr;
    size_t fBatchSize = 1;
    SolverMethod fSolverMethod;
    uint64_t fStep = 0;
    size_t fN;
    double fDx, fMass, fHbar;
    Cx fDt;
    int fBoundary;
    std::vector<Cx> fFTPsi, fIWPsi, fIWKPsi;
    kissfft<double> fft_N, inv_fft_N, inv_fft_2N;

    Split1D(size_t n, double dx, Cx dt, double mass, double hbar, int method, int boundary)
        : fN(n), fDx(dx), fMass(mass), fHbar(hbar), fDt(dt), fBoundary(boundary),
          fFTPsi(n), fIWPsi(2 * n), fIWKPsi(2 * n), fft_N(n, false), inv_fft_N(n, true),
          inv_fft_2N(2 * n, true)
    {
        fSolverMethod = method == 4 ? SolverMethod::SplittingMethodO4 : SolverMethod::SplittingMethodO2;
    }

    // src/SplittingMethod1D.cpp:84-89
    void ExpV(Cx* psi, double t)
    {
        Complex f = -I / fHbar * fDt * t;
        if (fVc) dev.MulExp(psi, f, (Cx const*)fVc, fN);
        else dev.MulExp(psi, f, fV, fN);
    }

    // src/SplittingMethod1D.cpp:40-81
    void ExpT(Cx* psi, double tt)
    {
        if (fBoundary == 0) {
            fft_N.transform(psi, fFTPsi.data());
            Real Dk = 2 * Pi / (fDx * fN);
            Real DT = QuSqr(Dk * fHbar) / (2 * fMass);
            Complex f = -I * DT * fDt * tt / fHbar;
            dev.MulExpK1D(fFTPsi.data(), f, fN);
            inv_fft_N.transform(fFTPsi.data(), psi);
            dev.Scale(psi, 1. / fN, fN);
        } else {
            Cx* a = fIWPsi.data();
            Cx* b = fIWKPsi.data();
            dev.Copy(a, psi, fN);
            dev.Set(a, 0, 0.);
            dev.Set(a, fN, 0.);
            dev.CopyReverseMinu(a + 2 * fN - 1, a + 1, fN - 1);
            inv_fft_2N.transform(a, b);
            double const Dp = fHbar * Pi / (fDx * fN);
            double const DT = Dp * Dp / (2 * fMass);
            dev.MulExpIdx2(b, -I * DT * fDt * tt / fHbar, fN);
            dev.Set(b, 0, 0.);
            dev.Set(b, fN, 0.);
            dev.CopyReverseMinu(b + 2 * fN - 1, b + 1, fN - 1);
            inv_fft_2N.transform(b, a);
            dev.Scale(psi, a, -1. / (2. * fN), fN);
        }
    }
};

// Split-operator 2D with the cached phase tables of SplittingMethod2D
// (src/SplittingMethod2D.cpp:38-121).  With a complex potential Vc the
// tables come from the reference's complex Exp (src/Device.h:104,
// src/DeviceTemplate.h:1311-1318) instead of the real one.
struct Split2D {
    Dev dev;
    Cx* fPsi = nullptr;
    size_t fBatchSize = 1;
    SolverMethod fSolverMethod;
    uint64_t fStep = 0;
    size_t fNx, fNy, fN;
    std::vector<Cx> ev05, evC1, evC2, et1, etD1, etD2, ft;
    Kiss2D fft, inv_fft;

    template <class VT>
    void exp_v(Cx* to, VT const* V, Cx f)
    {
        dev.Exp(to, V, f, fN);
    }

    template <class VT>
    Split2D(size_t nx, size_t ny, double dx, double dy, Cx dt, double mass, double hbar,
            int method, VT const* V)
        : fNx(nx), fNy(ny), fN(nx * ny), ft(nx * ny), fft(nx, ny, false), inv_fft(nx, ny, true)
    {
        fSolverMethod = method == 4 ? SolverMethod::SplittingMethodO4 : SolverMethod::SplittingMethodO2;
        if (method == 2) {
            ev05.resize(fN);
            ComplexType f = -I / hbar * dt * 0.5;
            exp_v(ev05.data(), V, f);
        } else {
            evC1.resize(fN);
            evC2.resize(fN);
            exp_v(evC1.data(), V, -I / hbar * dt * SplitingConstants<Real>::C1);
            exp_v(evC2.data(), V, -I / hbar * dt * SplitingConstants<Real>::C2);
        }
        RealType const Dkx = 2 * Pi / (dx * nx);
        RealType const Dky = 2 * Pi / (dy * ny);
        RealType const DTx = QuSqr(Dkx * hbar) / (2 * mass);
        RealType const DTy = QuSqr(Dky * hbar) / (2 * mass);
        ComplexType const alpha = -I * dt / hbar;
        if (method == 2) {
            et1.resize(fN);
            dev.SetOne(et1.data(), fN);
            dev.MulExpK2D(et1.data(), alpha, DTx, nx, DTy, ny);
        } else {
            etD1.resize(fN);
            etD2.resize(fN);
            dev.SetOne(etD1.data(), fN);
            dev.SetOne(etD2.data(), fN);
            dev.MulExpK2D(etD1.data(), alpha * SplitingConstants<Real>::D1, DTx, nx, DTy, ny);
            dev.MulExpK2D(etD2.data(), alpha * SplitingConstants<Real>::D2, DTx, nx, DTy, ny);
        }
    }
    using RealType = double;
    using ComplexType = Cx;

    void ExpV(Cx* psi, double tt)
    {
        if (tt == 0.5) dev.Mul(psi, ev05.data(), fN);
        else if (tt == SplitingConstants<Real>::C1) dev.Mul(psi, evC1.data(), fN);
        else dev.Mul(psi, evC2.data(), fN);
    }
    void ExpT(Cx* psi, double tt)
    {
        fft.Transform(psi, ft.data());
        if (tt == 1.0) dev.Mul(ft.data(), et1.data(), fN);
        else if (tt == SplitingConstants<Real>::D1) dev.Mul(ft.data(), etD1.data(), fN);
        else dev.Mul(ft.data(), etD2.data(), fN);
        inv_fft.Transform(ft.data(), psi);
        dev.Scale(psi, 1. / fN, fN);
    }
};

} // namespace

extern "C" {

// Every Device op of the CpuSeq device (src/DeviceTemplate.h) on the battery's
// fixed inputs (oracle/device_op_battery.h).
int ref_device_ops(size_t nx, size_t ny, size_t nz, const Cx* a, const Cx* b, const Cx* c,
                   const double* r, Cx* out)
{
    Dev dev;
    device_op_battery(dev, nx, ny, nz, a, b, c, r, out);
    return 0;
}

int ref_fft(int rank, const size_t* dims, int inverse, const Cx* in, Cx* out)
{
    try {
        if (rank == 1) {
            kissfft<double> f(dims[0], inverse != 0);
            f.transform(in, out);
        } else if (rank == 2) {
            Kiss2D f(dims[0], dims[1], inverse != 0);
            f.Transform(in, out);
        } else if (rank == 3) {
            Kiss3D f(dims[0], dims[1], dims[2], inverse != 0);
            f.Transform(in, out);
        } else {
            return -1;
        }
    } catch (...) {
        return -2;
    }
    return 0;
}

int ref_split1d_run(size_t n, double x0, double x1, double dt_re, double dt_im, double mass,
                    double hbar, int method, int boundary, const double* V, Cx* psi,
                    size_t nsteps, size_t batch)
{
    double dx = (x1 - x0) / n;
    Split1D s(n, dx, Cx(dt_re, dt_im), mass, hbar, method, boundary);
    std::vector<double> v(V, V + n);
    s.fV = v.data();
    s.fPsi = psi;
    s.fBatchSize = batch;
    for (size_t i = 0; i < nsteps; ++i) QuUpdatePsi(&s);
    return 0;
}

int ref_split2d_run(size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                    double dt_re, double dt_im, double mass, double hbar, int method,
                    const double* V, Cx* psi, size_t nsteps, size_t batch)
{
    std::vector<double> v(V, V + nx * ny);
    Split2D s(nx, ny, (x1 - x0) / nx, (y1 - y0) / ny, Cx(dt_re, dt_im), mass, hbar, method,
              v.data());
    s.fPsi = psi;
    s.fBatchSize = batch;
    for (size_t i = 0; i < nsteps; ++i) QuUpdatePsi(&s);
    return 0;
}

// QuEvolverImpl::Step's renormalisation (src/EvolverImpl.cpp:60-63):
// psi *= 1 / sqrt(Norm2() ), Norm2() = Device::Norm2 * dv (:51-54).
static void fnes_scale(Dev& dev, Cx* psi, size_t n, double dv)
{
    double scale = 1.0 / sqrt(dev.Norm2(psi, n) * dv);
    dev.Scale(psi, scale, n);
}

// The same evolutions with a complex potential (the reference's complex
// Exp / MulExp device ops) and optionally FNES: nsteps Step() calls, each
// `batch` QuUpdatePsi steps (src/EvolverImpl.cpp:56-68).
int ref_split1d_cv_run(size_t n, double x0, double x1, double dt_re, double dt_im, double mass,
                       double hbar, int method, int boundary, const Cx* V, Cx* psi,
                       size_t nsteps, size_t batch, int fnes)
{
    double dx = (x1 - x0) / n;
    Split1D s(n, dx, Cx(dt_re, dt_im), mass, hbar, method, boundary);
    std::vector<Cx> v(V, V + n);
    s.fVc = v.data();
    s.fPsi = psi;
    s.fBatchSize = batch;
    for (size_t i = 0; i < nsteps; ++i) {
        QuUpdatePsi(&s);
        if (fnes) fnes_scale(s.dev, psi, n, dx);
    }
    return 0;
}

int ref_split2d_cv_run(size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                       double dt_re, double dt_im, double mass, double hbar, int method,
                       const Cx* V, Cx* psi, size_t nsteps, size_t batch, int fnes)
{
    std::vector<Cx> v(V, V + nx * ny);
    const double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
    Split2D s(nx, ny, dx, dy, Cx(dt_re, dt_im), mass, hbar, method, (Cx const*)v.data());
    s.fPsi = psi;
    s.fBatchSize = batch;
    for (size_t i = 0; i < nsteps; ++i) {
        QuUpdatePsi(&s);
        if (fnes) fnes_scale(s.dev, psi, nx * ny, dx * dy);
    }
    return 0;
}

// Evolver observables through the reference's device ops.
// out[0] KinEn, out[1] PotEn, out[2] Norm2 (with dv), out[3] Xavg (1D only).
//   KinEn 1D: SplittingMethod1D::CalKinEn (src/SplittingMethod1D.cpp:94-128),
//             periodic (boundary 0) and infinite wall (1)
//   KinEn 2D: SplittingMethod2D::CalKinEn (src/SplittingMethod2D.cpp:124-141)
//   PotEn:    QuEvolverImpl::CalPotEn (src/EvolverImpl.cpp:45-49), real V
//   Norm2:    QuEvolverImpl::Norm2 (src/EvolverImpl.cpp:51-54)
//   Xavg:     QuEvolver1DImpl::Xavg (src/EvolverImpl1D.cpp:82-87)
int ref_observables(int dim, size_t nx, size_t ny, double x0, double x1, double y0, double y1,
                    double mass, double hbar, int boundary, const Cx* psi_in,
                    const double* V, double* out)
{
    try {
        Dev dev;
        const size_t n = dim == 1 ? nx : nx * ny;
        std::vector<Cx> psi(psi_in, psi_in + n);
        std::vector<double> v(V, V + n);
        const double dx = (x1 - x0) / nx, dy = dim == 1 ? 1.0 : (y1 - y0) / ny;
        const double dv = dx * dy;
        double kin = 0;
        if (dim == 1 && boundary == 0) {
            std::vector<Cx> ft(n);
            kissfft<double> f(n, false);
            f.transform(psi.data(), ft.data());
            Real Dk = 2 * Pi / (dx * n);
            Real DT = Dk * Dk / (2 * mass);
            kin = dev.Abs2K1D(ft.data(), DT, n) / dev.Norm2(ft.data(), n);
        } else if (dim == 1) {
            std::vector<Cx> a(2 * n), b(2 * n);
            kissfft<double> inv2n(2 * n, true);
            dev.Copy(a.data(), psi.data(), n);
            dev.Set(a.data(), 0, 0.);
            dev.Set(a.data(), n, 0.);
            dev.CopyReverseMinu(a.data() + 2 * n - 1, a.data() + 1, n - 1);
            inv2n.transform(a.data(), b.data());
            Real const Dk = Pi / (dx * n);
            Real const DT = QuSqr(Dk * hbar) / (2 * mass);
            kin = dev.Abs2Idx2(b.data(), n) * DT / dev.Norm2(b.data(), n);
        } else {
            std::vector<Cx> ft(n);
            Kiss2D f(nx, ny, false);
            f.Transform(psi.data(), ft.data());
            Real Dkx = 2 * Pi / (dx * nx);
            Real Dky = 2 * Pi / (dy * ny);
            Real DTx = QuSqr(Dkx * hbar) / (2 * mass);
            Real DTy = QuSqr(Dky * hbar) / (2 * mass);
            kin = dev.Abs2K2D(ft.data(), DTx, nx, DTy, ny) / dev.Norm2(ft.data(), n);
        }
        const double norm2 = dev.Norm2(psi.data(), n) * dv;
        Complex pot = dev.Abs2Mul(psi.data(), v.data(), n) * dv;
        out[0] = kin;
        out[1] = pot.real() / norm2;
        out[2] = norm2;
        out[3] = dim == 1 ? x0 + dx * (dev.Abs2Idx(psi.data(), n) / dev.Norm2(psi.data(), n)) : 0;
    } catch (...) {
        return -2;
    }
    return 0;
}

// Born series, src/Perturbation{2,3}DImpl.cpp init + Compute with the
// reference kernels of src/Perburbation.h.  ncalls Compute() calls of
// `order` iterations.  pre: 0 Identity, 1 Vellekoop, 2 Hao1, 3 Hao2.
int ref_born_run(int dim, const size_t* n, const double* lo, const double* hi, double en,
                 const double* dir, double mass, double hbar, int preconditional, int pre,
                 double slow, int order, const Cx* Vin, Cx* dpsi, Cx* v_eff, double* eps_inout,
                 double* delta_norm, int ncalls)
{
    try {
        Dev dev;
        size_t tot = n[0] * n[1] * (dim == 3 ? n[2] : 1);
        std::vector<Cx> V(Vin, Vin + tot), psi0(tot), psik(tot), tmp(tot), last(tot);
        double d[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) d[a] = (hi[a] - lo[a]) / n[a];
        // InitScatteringSolver: fK0 = sqrt(2 E m) / hbar; unit direction
        double k0 = QuCalMomentum(en, mass) / hbar;
        double nm2 = 0;
        for (int a = 0; a < dim; ++a) nm2 += dir[a] * dir[a];
        double nm = 1 / sqrt(nm2);
        double kk[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) kk[a] = k0 * (dir[a] * nm);
        if (dim == 2) {
            for_each_global_idx(lo[0], d[0], n[0], lo[1], d[1], n[1],
                                [&](size_t idx, Real x, Real y) {
                                    psi0[idx] = exp((kk[0] * x + kk[1] * y) * I);
                                });
        } else {
            for_each_global_idx(lo[0], d[0], n[0], lo[1], d[1], n[1], lo[2], d[2], n[2],
                                [&](size_t idx, Real x, Real y, Real z) {
                                    psi0[idx] = exp((kk[0] * x + kk[1] * y + kk[2] * z) * I);
                                });
        }
        double eps = *eps_inout;
        BornSerisePreconditioner P = pre == 0 ? BornSerisePreconditioner::Identity
                                   : pre == 1 ? BornSerisePreconditioner::Vellekoop
                                   : pre == 2 ? BornSerisePreconditioner::Hao1
                                              : BornSerisePreconditioner::Hao2;
        if (preconditional) {
            if (dim == 2)
                PerburbationUtility::GaussMAsbLayer2D(n[0], n[1], d[0], d[1], V.data(), hbar,
                                                      mass, en, 4.0);
            else
                PerburbationUtility::GaussMAsbLayer3D(n[0], n[1], n[2], d[0], d[1], d[2],
                                                      V.data(), hbar, mass, en, 4.0);
            PreconditionalBornSeries pbs;
            pbs.GetEpsilon(eps, P, tot, V.data(), &dev);
        }
        std::unique_ptr<FourierTransform2D> f2, i2;
        std::unique_ptr<FourierTransform3D> f3, i3;
        if (dim == 2) {
            f2.reset(new Kiss2D(n[0], n[1], false));
            i2.reset(new Kiss2D(n[0], n[1], true));
        } else {
            f3.reset(new Kiss3D(n[0], n[1], n[2], false));
            i3.reset(new Kiss3D(n[0], n[1], n[2], true));
        }
        double dn = 0;
        for (int c = 0; c < ncalls; ++c) {
            dev.Copy(last.data(), dpsi, tot);
            for (int it = 0; it < order; ++it) {
                if (preconditional) {
                    PreconditionalBornSeries pbs;
                    if (dim == 2)
                        pbs.Update2D(n[0], n[1], psi0.data(), dpsi, psik.data(), V.data(),
                                     tmp.data(), eps, en, mass, hbar, d[0], d[1], slow, P,
                                     f2.get(), i2.get(), &dev);
                    else
                        pbs.Update3D(n[0], n[1], n[2], psi0.data(), dpsi, psik.data(), V.data(),
                                     tmp.data(), eps, en, mass, hbar, d[0], d[1], d[2], slow, P,
                                     f3.get(), i3.get(), &dev);
                } else {
                    BornSeries bs;
                    if (dim == 2)
                        bs.Update2D(n[0], n[1], psi0.data(), dpsi, psik.data(), V.data(), eps,
                                    en, mass, hbar, d[0], d[1], f2.get(), i2.get(), &dev);
                    else
                        bs.Update3D(n[0], n[1], n[2], psi0.data(), dpsi, psik.data(), V.data(),
                                    eps, en, mass, hbar, d[0], d[1], d[2], f3.get(), i3.get(),
                                    &dev);
                }
            }
            Real num = dev.MinusNorm2(dpsi, last.data(), tot);
            Real den = dev.Norm2(dpsi, tot);
            dn = sqrt(num / den);
        }
        if (v_eff) std::memcpy(v_eff, V.data(), sizeof(Cx) * tot);
        *eps_inout = eps;
        if (delta_norm) *delta_norm = dn;
    } catch (...) {
        return -2;
    }
    return 0;
}

// ScatteringSolver3DImpl::ComputeXSection (src/ScatteringSolverImpl.cpp:318-361)
// through the reference Xsection3D reduction.
double ref_xsection3d(const size_t* n, const double* lo, const double* hi, double en,
                      double mass, double hbar, const Cx* psi0, const Cx* dpsi, const Cx* V,
                      double cosx, double cosy, double cosz)
{
    Dev dev;
    Real const nm = 1 / sqrt(cosx * cosx + cosy * cosy + cosz * cosz);
    cosx *= nm;
    cosy *= nm;
    cosz *= nm;
    double k0 = QuCalMomentum(en, mass) / hbar;
    double dx = (hi[0] - lo[0]) / n[0], dy = (hi[1] - lo[1]) / n[1], dz = (hi[2] - lo[2]) / n[2];
    Complex r = dev.Xsection3D(psi0, dpsi, V, k0 * cosx, k0 * cosy, k0 * cosz, lo[0], lo[1],
                               lo[2], dx, dy, dz, n[0], n[1], n[2]);
    return abs2(r * (2 * mass) / (hbar * hbar) * dx * dy * dz / (4 * Pi));
}

} // extern "C"
