"""Seeded synthetic workloads of the benchmark configurations.

Random draws come from a 64-bit Mersenne Twister (MT19937-64, the generator of
C++ std::mt19937_64) mapped to [0, 1) as (draw >> 11) * 2^-53, so the same
seed gives the same potentials from C++ or Python."""
import math

import numpy as np

SEED = 20261202


class MT19937_64:
    """MT19937-64 (Matsumoto & Nishimura, 64-bit variant)."""

    NN, MM = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UM, LM = 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed=5489):
        self.mt = [0] * self.NN
        self.mt[0] = seed & self.MASK
        for i in range(1, self.NN):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) \
                & self.MASK
        self.mti = self.NN

    def next(self):
        if self.mti >= self.NN:
            mt = self.mt
            for i in range(self.NN):
                x = (mt[i] & self.UM) | (mt[(i + 1) % self.NN] & self.LM)
                xa = x >> 1
                if x & 1:
                    xa ^= self.MATRIX_A
                mt[i] = mt[(i + self.MM) % self.NN] ^ xa
            self.mti = 0
        x = self.mt[self.mti]
        self.mti += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & self.MASK

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))


def grid(lo, hi, n):
    d = (hi - lo) / n
    return lo + d * np.arange(n), d


# ---------------------------------------------------------------- 1D (C1)
def c1_problem(n=2048):
    """1D tunnelling: x in [-100,100), psi0 = exp(i x - (x+30)^2/(4*5^2)),
    V = 0.6 exp(-x^2/2), hbar = m = 1, dt = 0.5, 1000 steps."""
    x, dx = grid(-100.0, 100.0, n)
    psi = np.exp(1j * x - (x + 30) ** 2 / (4 * 25.0))
    psi = psi / math.sqrt(float(np.sum(np.abs(psi) ** 2)) * dx)
    V = 0.6 * np.exp(-x ** 2 / 2)
    return dict(psi=psi, V=V, x0=-100.0, x1=100.0, dt=0.5, steps=1000, n=n)


# ---------------------------------------------------------------- 2D (C2/C3)
def bumps_2d(seed=SEED, count=16):
    g = MT19937_64(seed)
    out = []
    for _ in range(count):
        cx = g.uniform(-40, 40)
        cy = g.uniform(-40, 40)
        h = g.uniform(0, 1)
        w = g.uniform(2, 6)
        out.append((cx, cy, h, w))
    return out


def c2_problem(n=1024, L=150.0, seed=SEED, dt=0.2, steps=500, absorber=True):
    """2D packet through 16 Gaussian bumps on [-L, L)^2 with an imaginary
    absorbing layer -i 0.1 ((20 - d)/20)^2 for d < 20 (d = distance to the
    nearest edge)."""
    x, dx = grid(-L, L, n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    V = np.zeros((n, n))
    for cx, cy, h, w in bumps_2d(seed):
        V += h * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * w * w))
    W = np.zeros_like(V)
    if absorber:
        d = np.minimum(np.minimum(X + L, L - X), np.minimum(Y + L, L - Y))
        W = np.where(d < 20.0, 0.1 * ((20.0 - d) / 20.0) ** 2, 0.0)
    psi = np.exp(1j * 0.8 * X - ((X + 60) ** 2 + Y ** 2) / (4 * 64.0))
    psi = psi / math.sqrt(float(np.sum(np.abs(psi) ** 2)) * dx * dx)
    return dict(psi=psi, V=V - 1j * W, x0=-L, x1=L, dt=dt, steps=steps, n=n)


def c3_problem(n=2048):
    return c2_problem(n=n, dt=0.1, steps=1000)


# ---------------------------------------------------------------- 3D (C4)
def gaussians_3d(seed=SEED, count=8):
    g = MT19937_64(seed + 3)
    out = []
    for _ in range(count):
        a = g.uniform(0, 0.3)
        w = g.uniform(1.5, 4)
        c = (g.uniform(-10, 10), g.uniform(-10, 10), g.uniform(-10, 10))
        out.append((a, w, c))
    return out


def c4_potential(n=128, L=32.0, seed=SEED):
    x, dx = grid(-L, L, n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    V = np.zeros((n, n, n))
    for a, w, c in gaussians_3d(seed):
        V += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * w * w))
    return V.astype(np.complex128)


def c4_energies(count=32, seed=SEED):
    g = MT19937_64(seed + 4)
    ks = [g.uniform(1, 3) for _ in range(count)]
    return [0.5 * k * k for k in ks]
