"""Synthetic SCFT workloads of the benchmark configurations.

The reference reads initial fields from a MAT file (gridInitType = File,
src/Grid.cc:694-711) and ships no gyroid / cylinder templates (its
GYROID_PATTERN is declared but not implemented, src/Helper.cc:28-52), so the
templates below are defined here.  Noise is Philox4x32-10 (numpy's
implementation) + normal sampling, seeded per configuration and per field
(stream 0 = wA, stream 1 = wB), so CPU and GPU runs see identical bits.

Ms follows the fork's convention (src/Model_AB.cc:100-111): LA = round(f Ns),
sA = LA + 1, sB = Ns - LA + 1, ds = 1/Ns.
"""
import math
from dataclasses import dataclass, field

import numpy as np


def noise(seed, stream, shape, sigma):
    g = np.random.Generator(np.random.Philox(np.random.SeedSequence([int(seed), int(stream)])))
    return sigma * g.standard_normal(shape)


def ms_from(f, Ns):
    LA = int(math.floor(f * Ns + 0.5))
    return LA + 1, Ns - LA + 1


def coords(shape, lengths):
    """Physical coordinates x_i = i * l / L of the C-order grid."""
    axes = [np.arange(n) * (l / n) for n, l in zip(shape, lengths)]
    return np.meshgrid(*axes, indexing="ij")


def gyroid(shape, lengths):
    X, Y, Z = [2 * np.pi * c / l for c, l in zip(coords(shape, lengths), lengths)]
    return np.sin(X) * np.cos(Y) + np.sin(Y) * np.cos(Z) + np.sin(Z) * np.cos(X)


def hex_cylinders(shape, lengths):
    """Three first-shell plane waves in the x-y plane (hexagonal cylinder template)."""
    X, Y, _ = coords(shape, lengths)
    kx, ky = 2 * np.pi / lengths[0], 2 * np.pi / lengths[1]
    return (np.cos(kx * X) + np.cos(0.5 * kx * X + ky * Y) + np.cos(0.5 * kx * X - ky * Y)) / 1.5


def lamellar(shape, lengths):
    X = coords(shape, lengths)[0]
    return np.cos(2 * np.pi * X / lengths[0])


@dataclass
class Workload:
    name: str
    shape: tuple
    cell: tuple          # (crystal, a, b, c)
    Ns: int
    f: float
    chiN: float
    wA: np.ndarray = field(repr=False)
    wB: np.ndarray = field(repr=False)
    lam: tuple = (0.1, 0.1)
    n_anderson: int = 20

    @property
    def Ms(self):
        return ms_from(self.f, self.Ns)

    def model_kwargs(self):
        return dict(cell=self.cell, Ms=self.Ms, chiN=self.chiN, lam=self.lam, n_anderson=self.n_anderson)


def cfg1(L=32):
    """AB melt f=0.3, chiN=20, cubic a=4.0, Ns=100, noisy homogeneous start."""
    shape = (L, L, L)
    f, chiN = 0.3, 20.0
    wA = chiN * (1 - f) + noise(1001, 0, shape, 0.5)
    wB = chiN * f + noise(1001, 1, shape, 0.5)
    return Workload("cfg1", shape, ("CUBIC", 4.0, 4.0, 4.0), 100, f, chiN, wA, wB)


def cfg2(L=128, Ns=200, replica=0):
    """AB melt f=0.35, chiN=18, cubic a=4.2, 128^3, Ns=200, gyroid template.

    replica r > 0 draws the noise from streams 2r (wA) and 2r + 1 (wB) of the
    same seed: independent copies of the system for batched solves."""
    shape = (L, L, L)
    f, chiN, a = 0.35, 18.0, 4.2
    g = gyroid(shape, (a, a, a))
    wA = chiN * (1 - f) + 2 * g + noise(1002, 2 * replica, shape, 0.2)
    wB = chiN * f - 2 * g + noise(1002, 2 * replica + 1, shape, 0.2)
    return Workload("cfg2", shape, ("CUBIC", a, a, a), Ns, f, chiN, wA, wB)


def cfg4(L=64):
    """AB melt f=0.3, chiN=20, orthorhombic 3.8 x 4.0 x 4.3, 64^3, Ns=200, hex cylinders."""
    shape = (L, L, L)
    f, chiN = 0.3, 20.0
    lengths = (3.8, 4.0, 4.3)
    h = hex_cylinders(shape, lengths)
    wA = chiN * (1 - f) + 2 * h + noise(1004, 0, shape, 0.1)
    wB = chiN * f - 2 * h + noise(1004, 1, shape, 0.1)
    return Workload("cfg4", shape, ("ORTHORHOMBIC",) + lengths, 200, f, chiN, wA, wB)


@dataclass
class StringWorkload:
    name: str
    shape: tuple
    cell: tuple
    Ns: int
    f: float
    chiN: float
    wA: np.ndarray = field(repr=False)   # (m, Lx, Ly, Lz)
    wB: np.ndarray = field(repr=False)
    lam: tuple = (0.1, 0.1, 1.0)
    respline_every: int = 10

    @property
    def Ms(self):
        return ms_from(self.f, self.Ns)

    @property
    def m(self):
        return self.wA.shape[0]

    def string_kwargs(self):
        return dict(cell=self.cell, Ms=self.Ms, chiN=self.chiN, lam=self.lam)


def cfg5a(m=32, L=64, Ns=150):
    """String of m images, AB melt f=0.33, chiN=17.5, 64^3, Ns=150: ends are
    lamellar and hexagonal-cylinder templates, the inner images their linear
    interpolation, plus noise sigma 0.05 (seed 1005, streams 0 / 1 per field,
    2k / 2k+1 per image)."""
    shape = (L, L, L)
    f, chiN, a = 0.33, 17.5, 4.0
    lengths = (a, a, a)
    ends = [lamellar(shape, lengths), hex_cylinders(shape, lengths)]
    t = np.linspace(0.0, 1.0, m)
    wA = np.empty((m,) + shape)
    wB = np.empty((m,) + shape)
    for k in range(m):
        g = (1 - t[k]) * ends[0] + t[k] * ends[1]
        wA[k] = chiN * (1 - f) + 2 * g + noise(1005, 2 * k, shape, 0.05)
        wB[k] = chiN * f - 2 * g + noise(1005, 2 * k + 1, shape, 0.05)
    return StringWorkload("cfg5a", shape, ("CUBIC", a, a, a), Ns, f, chiN, wA, wB)


def cfg5b(n=24, sizes=(32, 48, 64, 96), Ns_choices=(100, 150, 200)):
    """Ragged scan: n independent solves with grids, Ns, f ~ U[.25, .45] and
    chiN ~ U[14, 22] drawn from seed 1005 (stream 99), each started from a
    gyroid template plus noise sigma 0.2 on a cubic cell a = 4.0."""
    g = np.random.Generator(np.random.Philox(np.random.SeedSequence([1005, 99])))
    out = []
    for k in range(n):
        L = int(sizes[g.integers(len(sizes))])
        Ns = int(Ns_choices[g.integers(len(Ns_choices))])
        f = float(g.uniform(0.25, 0.45))
        chiN = float(g.uniform(14.0, 22.0))
        shape, a = (L, L, L), 4.0
        gy = gyroid(shape, (a, a, a))
        wA = chiN * (1 - f) + 2 * gy + noise(1005, 100 + 2 * k, shape, 0.2)
        wB = chiN * f - 2 * gy + noise(1005, 101 + 2 * k, shape, 0.2)
        out.append(Workload(f"cfg5b-{k}", shape, ("CUBIC", a, a, a), Ns, f, chiN, wA, wB))
    return out


def contour_steps_per_iteration(Ns):
    """Counted contour steps per SCFT iteration: qA, qB, qBc, qAc = 2 Ns."""
    return 2 * Ns


def algorithmic_bytes_per_step(shape, p):
    """(13 M + 12 Mh) p: each FFT one HBM round trip, pointwise factors fused."""
    Lx, Ly, Lz = shape
    M = Lx * Ly * Lz
    Mh = Lx * Ly * (Lz // 2 + 1)
    return (13 * M + 12 * Mh) * p


def etdrk4_bytes_per_step(shape, p):
    """(5 M + 64 Mh) p: the bytes the ETDRK4 dataflow moves per contour step.
    Four stages, each an x-pencil pass and a yz-plane pass that read and write
    the spectrum (8 Mh p) and read w (M p); the slice store (M p); the
    carried spectra v, acc, t (11 complex reads / writes, 22 Mh p) and the
    coefficient tables E, E2, Q, f1, f2, f3 (10 reads, 10 Mh p)."""
    Lx, Ly, Lz = shape
    M = Lx * Ly * Lz
    Mh = Lx * Ly * (Lz // 2 + 1)
    return (5 * M + 64 * Mh) * p


def adjoint_bytes_per_lockstep_step(shape, p):
    """Extra bytes of the fused stress adjoint per lockstep step (two chains):
    each chain stores three full spectra per step in the first half of the
    contour and reads its partner's three in the second half, 3 Mh complex
    (6 Mh p bytes) per chain-step on average."""
    Lx, Ly, Lz = shape
    Mh = Lx * Ly * (Lz // 2 + 1)
    return 2 * 6 * Mh * p
