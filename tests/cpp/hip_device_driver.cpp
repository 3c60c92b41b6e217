// hip_device_driver.cpp -- the reference's solver-side call sequences run on
// HipDevice through the Device / FourierTransform plug-in interfaces, the way
// the reference's solvers would call them once Device::Create(GPU_HIP)
// returns a HipDevice.  Reads inputs from and writes results to raw binary
// files; the parity test compares them with the oracle.
//
//   split2d <nx> <ny> <x0> <x1> <y0> <y1> <dt> <steps> <in.bin> <out.bin>
//     SplittingMethod2D O2 with cached tables (src/SplittingMethod2D.cpp:38-121),
//     V real (the reference evolver), input: psi (nx*ny c128) then V (nx*ny f64).
//   born3d <n> <L> <E> <eps> <slow> <order> <in.bin> <out.bin>
//     PreconditionalBornSeries::Update3D with Vellekoop (src/Perburbation.h:350-531)
//     from precomputed psi0 and V_eff; input: psi0 then V_eff (n^3 c128 each).
//   ops <nx> <ny> <nz> <in.bin> <out.bin>
//     every Device op once on fixed inputs (oracle/device_op_battery.h);
//     input: a, b, c (n c128 each) then r (n f64).
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hip_device.h"
#include "../../oracle/device_op_battery.h"

using namespace qsim_host;
using cd = std::complex<double>;

static std::vector<char> slurp(const char *path)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> b(n);
    if (std::fread(b.data(), 1, n, f) != (size_t)n) throw std::runtime_error("short read");
    std::fclose(f);
    return b;
}

static void spit(const char *path, const void *p, size_t bytes)
{
    FILE *f = std::fopen(path, "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("write failed");
    std::fclose(f);
}

static int split2d(char **a)
{
    size_t nx = std::atol(a[0]), ny = std::atol(a[1]);
    double x0 = std::atof(a[2]), x1 = std::atof(a[3]), y0 = std::atof(a[4]), y1 = std::atof(a[5]);
    double dt = std::atof(a[6]);
    int steps = std::atoi(a[7]);
    size_t N = nx * ny;
    auto in = slurp(a[8]);
    HipDevice dev;
    cd *psi = dev.Alloc<cd>(N), *ft = dev.Alloc<cd>(N), *ev = dev.Alloc<cd>(N), *et = dev.Alloc<cd>(N);
    double *V = dev.Alloc<double>(N);
    dev.ToDevice(psi, (const cd *)in.data(), N);
    dev.ToDevice(V, (const double *)(in.data() + N * sizeof(cd)), N);
    const double hbar = 1, mass = 1, dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
    // InitExpV / InitExpT
    dev.Exp(ev, V, cd(-0.0, -1.0) / hbar * dt * 0.5, N);
    double Dkx = 2 * M_PI / (dx * nx), Dky = 2 * M_PI / (dy * ny);
    double DTx = (Dkx * hbar) * (Dkx * hbar) / (2 * mass), DTy = (Dky * hbar) * (Dky * hbar) / (2 * mass);
    dev.SetOne(et, N);
    dev.MulExpK2D(et, cd(-0.0, -1.0) * dt / hbar, DTx, nx, DTy, ny);
    HipFourierTransform2D fft(nx, ny, false), ifft(nx, ny, true);
    for (int s = 0; s < steps; ++s) {
        dev.Mul(psi, ev, N);
        fft.Transform(psi, ft);
        dev.Mul(ft, et, N);
        ifft.Transform(ft, psi);
        dev.Scale(psi, 1. / N, N);
        dev.Mul(psi, ev, N);
    }
    std::vector<cd> out(N);
    dev.ToHost(out.data(), psi, N);
    spit(a[9], out.data(), N * sizeof(cd));
    std::printf("norm2*dv %.17g\n", dev.Norm2(psi, N) * dx * dy);
    for (void *p : {(void *)psi, (void *)ft, (void *)ev, (void *)et, (void *)V}) dev.Free(p);
    return 0;
}

static int born3d(char **a)
{
    size_t n = std::atol(a[0]);
    double L = std::atof(a[1]), E = std::atof(a[2]), eps = std::atof(a[3]), slow = std::atof(a[4]);
    int order = std::atoi(a[5]);
    size_t N = n * n * n;
    auto in = slurp(a[6]);
    HipDevice dev;
    cd *psi0 = dev.Alloc<cd>(N), *v = dev.Alloc<cd>(N), *dp = dev.Alloc<cd>(N),
       *psik = dev.Alloc<cd>(N), *tmp = dev.Alloc<cd>(N);
    dev.ToDevice(psi0, (const cd *)in.data(), N);
    dev.ToDevice(v, (const cd *)(in.data() + N * sizeof(cd)), N);
    dev.SetZero(dp, N);
    HipFourierTransform3D fft(n, n, n, false), ifft(n, n, n, true);
    const double d = 2 * L / n, Dp = 2 * M_PI / (d * n), a1 = Dp * Dp / 2;
    for (int it = 0; it < order; ++it) {
        dev.Add3(tmp, dp, psi0, v, eps, N);
        fft.Transform(tmp, psik);
        dev.G03D(psik, E, eps, a1, a1, a1, n, n, n);
        ifft.Transform(psik, tmp);
        dev.Scale(tmp, 1. / N, N);
        dev.Vellekoop(dp, tmp, v, slow, eps, N);
    }
    std::vector<cd> out(N);
    dev.ToHost(out.data(), dp, N);
    spit(a[7], out.data(), N * sizeof(cd));
    cd r = dev.Xsection3D(psi0, dp, v, 0, 0, std::sqrt(2 * E), -L, -L, -L, d, d, d, n, n, n);
    std::printf("xsec_sum %.17g %.17g\n", r.real(), r.imag());
    for (void *p : {(void *)psi0, (void *)v, (void *)dp, (void *)psik, (void *)tmp}) dev.Free(p);
    return 0;
}

static int ops(char **a)
{
    size_t nx = std::atol(a[0]), ny = std::atol(a[1]), nz = std::atol(a[2]);
    size_t N = nx * ny * nz;
    auto in = slurp(a[3]);
    if (in.size() != N * (3 * sizeof(cd) + sizeof(double))) throw std::runtime_error("bad input size");
    const cd *pa = (const cd *)in.data();
    const double *pr = (const double *)(in.data() + 3 * N * sizeof(cd));
    std::vector<cd> out(N * kBatteryArrays + kBatteryScalars);
    HipDevice dev;
    device_op_battery(dev, nx, ny, nz, pa, pa + N, pa + 2 * N, pr, out.data());
    spit(a[4], out.data(), out.size() * sizeof(cd));
    return 0;
}

int main(int argc, char **argv)
{
    try {
        if (argc >= 7 && std::strcmp(argv[1], "ops") == 0) return ops(argv + 2);
        if (argc >= 12 && std::strcmp(argv[1], "split2d") == 0) return split2d(argv + 2);
        if (argc >= 10 && std::strcmp(argv[1], "born3d") == 0) return born3d(argv + 2);
        std::fprintf(stderr, "usage: split2d ... | born3d ...\n");
        return 2;
    } catch (std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
