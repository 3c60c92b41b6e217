"""Writes tests/golden/ref_vectors.npz: outputs of the compiled reference
(oracle/_ref/libqsim_ref.so, built from the reference's own kissfft.hh,
DeviceTemplate.h, Perburbation.h and SplittingUtils sources) on small seeded
inputs.  Run in a tree where the reference is present:

    make -C oracle ref && python tests/golden/make_golden.py

The fixture holds inputs where they are not a simple formula, and outputs."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402


def ramp(shape):
    """x_n = n, the reference FFT test input (applications/QuSimTest/TestFFT.cpp:27-30)."""
    return np.arange(int(np.prod(shape)), dtype=np.float64).reshape(shape).astype(np.complex128)


def main():
    out = {}
    # FFT: the reference test sizes, forward and inverse
    for name, shape in (("fft1234", (1234,)), ("fft123x234", (123, 234)),
                        ("fft12x23x34", (12, 23, 34)), ("fft64x48", (64, 48))):
        x = ramp(shape)
        f, i = O.ref_fft(x, False), O.ref_fft(x, True)
        if f.size > 4096:  # keep the fixture small: a fixed sample + sums
            idx = np.random.default_rng(7).choice(f.size, 1024, replace=False)
            out[name + "_idx"] = idx
            out[name + "_fwd_sample"] = f.ravel()[idx]
            out[name + "_inv_sample"] = i.ravel()[idx]
            out[name + "_sums"] = np.array([f.sum(), (np.abs(f) ** 2).sum(), i.sum()])
        else:
            out[name + "_fwd"] = f
            out[name + "_inv"] = i
    # 1D split: the C1 configuration at full size, O2 and O4, and the wall
    n = 2048
    x = -100 + 200.0 / n * np.arange(n)
    psi = np.exp(1j * x - (x + 30) ** 2 / 100.0)
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * (200.0 / n))
    V = 0.6 * np.exp(-x ** 2 / 2)
    out["c1_psi0"] = psi
    out["c1_o2_1000"] = O.ref_split1d(psi, V, -100.0, 100.0, 0.5, 1000, method=2)
    out["c1_o4_200"] = O.ref_split1d(psi, V, -100.0, 100.0, 0.5, 200, method=4)
    out["c1_wall_o2_200"] = O.ref_split1d(psi, V, -100.0, 100.0, 0.5, 200, method=2,
                                          boundary=1)
    out["c1_cdt_o2_100"] = O.ref_split1d(psi, V, -100.0, 100.0, 0.5 - 0.05j, 100, method=2)
    # 2D split: 64 x 48, O2 and O4
    nx, ny = 64, 48
    gx = -30 + 60.0 / nx * np.arange(nx)
    gy = -20 + 45.0 / ny * np.arange(ny)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    p2 = np.exp(0.8j * X - ((X + 5) ** 2 + Y ** 2) / 36.0)
    V2 = 0.5 * np.exp(-(X ** 2 + Y ** 2) / 8)
    out["s2_psi0"] = p2
    out["s2_V"] = V2
    out["s2_o2_50"] = O.ref_split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 50, method=2)
    out["s2_o4_20"] = O.ref_split2d(p2, V2, -30.0, 30.0, -20.0, 25.0, 0.1, 20, method=4)
    # Complex potentials (the imaginary absorbing layer of the 2D workloads)
    # through the reference's complex Exp / MulExp device ops, and the FNES
    # renormalisation of Step().  1D: the C1 grid with -i 0.1 (d/20)^2 within
    # d < 20 of either edge.
    d1 = np.minimum(x + 100.0, 100.0 - x)
    V1c = V - 1j * np.where(d1 < 20.0, 0.1 * (d1 / 20.0) ** 2, 0.0)
    out["c1cv_V"] = V1c
    out["c1cv_o2_1000"] = O.ref_split1d_cv(psi, V1c, -100.0, 100.0, 0.5, 1000, method=2)
    out["c1cv_o2_150"] = O.ref_split1d_cv(psi, V1c, -100.0, 100.0, 0.5, 150, method=2)
    out["c1cv_o4_150"] = O.ref_split1d_cv(psi, V1c, -100.0, 100.0, 0.5, 150, method=4)
    out["c1cv_fnes_o2_120x2"] = O.ref_split1d_cv(psi, V1c, -100.0, 100.0, 0.5, 120, method=2,
                                                 batch=2, fnes=True)
    # 2D: 64 x 48 with -i 0.2 (d/8)^2 within d < 8 of the nearest edge
    dd = np.minimum.reduce([X + 30.0, 30.0 - X, Y + 20.0, 25.0 - Y])
    V2c = V2 - 1j * np.where(dd < 8.0, 0.2 * (dd / 8.0) ** 2, 0.0)
    out["s2cv_V"] = V2c
    out["s2cv_o2_80"] = O.ref_split2d_cv(p2, V2c, -30.0, 30.0, -20.0, 25.0, 0.1, 80, method=2)
    out["s2cv_o4_30"] = O.ref_split2d_cv(p2, V2c, -30.0, 30.0, -20.0, 25.0, 0.1, 30, method=4)
    out["s2cv_fnes_o2_40x2"] = O.ref_split2d_cv(p2, V2c, -30.0, 30.0, -20.0, 25.0, 0.1, 40,
                                                method=2, batch=2, fnes=True)
    # Evolver observables (KinEn, PotEn, Norm2, Xavg) through the reference's
    # Abs2K1D / Abs2K2D / Abs2Idx2 / Abs2Mul / Abs2Idx / Norm2 device ops, on
    # states that are not Gaussian any more
    s1 = out["c1cv_o2_150"]
    out["obs_c1_periodic"] = np.array(O.ref_observables(s1, V, -100.0, 100.0, boundary=0))
    out["obs_c1_wall"] = np.array(O.ref_observables(out["c1_wall_o2_200"], V, -100.0, 100.0,
                                                    boundary=1))
    out["obs_c1_hbar_mass"] = np.array(O.ref_observables(s1, V, -100.0, 100.0, mass=1.7,
                                                         hbar=0.8, boundary=1))
    out["obs_s2"] = np.array(O.ref_observables(out["s2cv_o2_80"], V2, -30.0, 30.0, -20.0, 25.0,
                                               mass=1.3, hbar=0.9))
    # Born 2D (48 x 40) and 3D (12 x 10 x 14), each preconditioner
    lo2, hi2 = (-60.0, -50.0), (60.0, 50.0)
    gx = lo2[0] + 120.0 / 48 * np.arange(48)
    gy = lo2[1] + 100.0 / 40 * np.arange(40)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    Vb2 = 0.3 * np.exp(-(X ** 2 + Y ** 2) / 4) + 0.05j * np.exp(-(X ** 2 + Y ** 2) / 9)
    out["b2_V"] = Vb2
    for tag, pc, pre, slow in (("naive", 0, 0, 1.0), ("ident", 1, 0, 0.6), ("vell", 1, 1, 0.7),
                               ("hao1", 1, 2, 0.7), ("hao2", 1, 3, 0.5)):
        dp, veff, eps, dn = O.ref_born(Vb2, lo2, hi2, 2.0, (1.0, 0.3), order=3, ncalls=2,
                                       preconditional=pc, preconditioner=pre, slow=slow,
                                       epsilon=0.01)
        out["b2_%s_dpsi" % tag] = dp
        out["b2_%s_scalars" % tag] = np.array([eps, dn])
        if tag == "vell":
            out["b2_vell_veff"] = veff
    lo3, hi3 = (-6.0, -5.0, -7.0), (6.0, 5.0, 7.0)
    g3 = [lo3[i] + (hi3[i] - lo3[i]) / s * np.arange(s) for i, s in enumerate((12, 10, 14))]
    X, Y, Z = np.meshgrid(*g3, indexing="ij")
    Vb3 = (0.2 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 4)).astype(np.complex128)
    out["b3_V"] = Vb3
    for tag, pc, pre, slow in (("naive", 0, 0, 1.0), ("vell", 1, 1, 0.5), ("hao2", 1, 3, 0.5)):
        dp, veff, eps, dn = O.ref_born(Vb3, lo3, hi3, 4.0, (0.0, 0.2, 1.0), order=3, ncalls=1,
                                       preconditional=pc, preconditioner=pre, slow=slow,
                                       epsilon=0.01)
        out["b3_%s_dpsi" % tag] = dp
        out["b3_%s_scalars" % tag] = np.array([eps, dn])
        if tag == "vell":
            out["b3_vell_veff"] = veff
            psi0 = O.plane_wave(Vb3.shape, lo3, hi3, 4.0, (0.0, 0.2, 1.0))
            out["b3_vell_xsec"] = np.array([O.ref_xsection3d(psi0, dp, veff, lo3, hi3, 4.0,
                                                             0.3, 0.1, 0.9)])
    # Every Device op of the CpuSeq device on fixed inputs (n = 3 x 4 x 5),
    # the battery of oracle/device_op_battery.h
    dims = (3, 4, 5)
    m = np.arange(60, dtype=np.float64)
    oa = np.sin(0.37 * m + 0.1) + 1j * np.cos(0.91 * m)
    ob = np.cos(0.23 * m - 0.4) + 0.5j * np.sin(0.71 * m + 0.2)
    oc = 0.3 * np.exp(-((m - 30) / 12) ** 2) + 0.05j * np.cos(0.17 * m)
    orr = 0.5 + 0.4 * np.sin(0.53 * m) ** 2
    arrays, scalars = O.ref_device_ops(dims, oa, ob, oc, orr)
    out.update(ops_dims=np.array(dims), ops_a=oa, ops_b=ob, ops_c=oc, ops_r=orr,
               ops_arrays=arrays, ops_scalars=scalars)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
