"""Seeded synthetic systems for the batched VMC walker-move path.

A system is a periodic orthorhombic cell with ions, two spin groups of
electrons, one real periodic cubic B-spline orbital table and one-/two-body
B-spline Jastrow functors, as used by QMCPACK's SplineR2R, J1OrbitalSoA and
TwoBodyJastrow<BsplineFunctor> (reference: src/QMCWaveFunctions/BsplineFactory/
SplineR2R.h:35-233, src/QMCWaveFunctions/Jastrow/BsplineFunctor.h:104-133).

Orbital coefficients are the unique periodic table u[Mx][My][Mz][norb]
(orbital fastest), drawn U[-1, 1] plus 0.5*cos(G_o . x + phi_o) with a seeded
reciprocal-lattice vector G_o and phase phi_o per orbital.  The library wraps
them into the padded [M+3]^3 layout itself (einspline/bspline_create.c:1027-1034).

Jastrow parameters follow P_i = 0.5*exp(-i/2); the cutoff is the Wigner-Seitz
radius L/2 of the cubic cell (RadialJastrowBuilder.cpp:215); cusps are
-0.25 (like spins) and -0.5 (unlike spins) from RadialJastrowBuilder.cpp:200-214,
and 0 for J1.
"""
from __future__ import annotations

import dataclasses
import numpy as np

# Workload definitions.  Sizes are the ones the benchmark and parity tests run.
WORKLOADS = {
    # plain method, small: CPU-runnable (16 walkers x 128 electrons, one crowd, delay rank 1)
    "small": dict(L=10.0, n_ion=16, n_up=64, n_dn=64, norb=64, M=48, nw=16, walkers_per_crowd=16,
            delay_rank=1, tau=0.3, sweeps=10, seed=20240),
    # plain method at production size (1024 walkers x 512 electrons, delay rank 32)
    "production": dict(L=18.0, n_ion=128, n_up=256, n_dn=256, norb=256, M=64, nw=1024, walkers_per_crowd=1,
            delay_rank=32, tau=0.3, sweeps=50, recompute_period=64, seed=20241),
}


@dataclasses.dataclass
class System:
    L: np.ndarray            # [3] cell lengths (orthorhombic)
    ions: np.ndarray         # [n_ion][3]
    n_up: int
    n_dn: int
    M: tuple                 # grid points per dimension
    norb: int
    coefs: np.ndarray | None  # [Mx][My][Mz][norb] float64 (unique periodic table)
    j1_params: np.ndarray
    j1_rcut: float
    j1_cusp: float
    j2_uu: np.ndarray
    j2_ud: np.ndarray
    j2_rcut: float
    j2_cusp_uu: float
    j2_cusp_ud: float
    delay_rank: int
    # complex (C2C) orbitals: coefs [Mx][My][Mz][norb][2] (real, imaginary) and the Cartesian k-point
    # of each orbital; psi_j(r) = exp(-i k_j . r) u_j(r) (SplineC2C.cpp:200-279)
    kpoints: np.ndarray | None = None

    @property
    def N(self) -> int:
        return self.n_up + self.n_dn


def jastrow_params(n: int = 8) -> np.ndarray:
    return 0.5 * np.exp(-0.5 * np.arange(n, dtype=np.float64))


def make_coefs(M: int, norb: int, rng: np.random.Generator, chunk: int = 32) -> np.ndarray:
    """u[M][M][M][norb] = U[-1,1] + 0.5*cos(2*pi*(n . idx)/M + phi), seeded."""
    coefs = rng.uniform(-1.0, 1.0, size=(M, M, M, norb))
    nvec = rng.integers(-2, 3, size=(norb, 3))
    phi = rng.uniform(0.0, 2.0 * np.pi, size=norb)
    i = np.arange(M)
    for o0 in range(0, norb, chunk):
        o1 = min(norb, o0 + chunk)
        # phase index m = (n1 i + n2 j + n3 k) mod M, exact in integers
        m = (nvec[o0:o1, 0, None, None, None] * i[None, :, None, None]
             + nvec[o0:o1, 1, None, None, None] * i[None, None, :, None]
             + nvec[o0:o1, 2, None, None, None] * i[None, None, None, :]) % M
        table = 0.5 * np.cos(2.0 * np.pi * np.arange(M)[None, :] / M + phi[o0:o1, None])  # [o][M]
        pw = np.take_along_axis(table.reshape(o1 - o0, M), m.reshape(o1 - o0, -1), axis=1)
        coefs[..., o0:o1] += pw.reshape(o1 - o0, M, M, M).transpose(1, 2, 3, 0)
    return np.ascontiguousarray(coefs)


def make_system(L: float, n_ion: int, n_up: int, n_dn: int, norb: int, M: int, delay_rank: int,
                seed: int, with_coefs: bool = True, n_jastrow: int = 8, **_ignored) -> System:
    rng = np.random.default_rng(seed)
    ions = rng.uniform(0.0, L, size=(n_ion, 3))
    coefs = make_coefs(M, norb, rng) if (with_coefs and norb > 0) else None
    p = jastrow_params(n_jastrow)
    rcut = 0.5 * L
    return System(L=np.array([L, L, L], dtype=np.float64), ions=ions, n_up=n_up, n_dn=n_dn,
                  M=(M, M, M), norb=norb, coefs=coefs,
                  j1_params=p.copy(), j1_rcut=rcut, j1_cusp=0.0,
                  j2_uu=p.copy(), j2_ud=p.copy(), j2_rcut=rcut, j2_cusp_uu=-0.25, j2_cusp_ud=-0.5,
                  delay_rank=delay_rank)


def make_complex_system(L: float, n_ion: int, n_up: int, n_dn: int, norb: int, M: int, delay_rank: int,
                        seed: int, n_jastrow: int = 8, **_ignored) -> System:
    """Complex (C2C) orbitals at a seeded twist: u_j = a_j + i b_j with a_j, b_j drawn as the real
    tables, and k_j = 2 pi (t + n_j) / L for the twist t in [-1/2, 1/2)^3 and the integer vectors n_j
    of smallest |t + n_j| (distinct per orbital)."""
    s = make_system(L, n_ion, n_up, n_dn, norb, M, delay_rank, seed, with_coefs=False, n_jastrow=n_jastrow)
    rng = np.random.default_rng(seed + 104729)
    re = make_coefs(M, norb, rng)
    im = make_coefs(M, norb, rng)
    coefs = np.ascontiguousarray(np.stack([re, im], axis=-1))
    twist = rng.uniform(-0.5, 0.5, size=3)
    r = int(np.ceil(np.cbrt(norb))) + 1
    n = np.array([(a, b, c) for a in range(-r, r + 1) for b in range(-r, r + 1) for c in range(-r, r + 1)],
                 dtype=np.float64)
    kv = twist[None, :] + n
    order = np.lexsort((kv[:, 2], kv[:, 1], kv[:, 0], np.round(np.sum(kv * kv, axis=1), 12)))
    kpoints = 2.0 * np.pi * kv[order[:norb]] / np.asarray(s.L)[None, :]
    return dataclasses.replace(s, coefs=coefs, kpoints=np.ascontiguousarray(kpoints))


# Uneven batching: 1024 walkers over four systems of 128 / 256 / 512 / 1024 electrons, each in its
# own cubic cell at the electron density of the 512-electron system (L = 18 (N / 512)^(1/3)) with one
# ion per four electrons and its own table (48^3 / 64^3 / 64^3 / 80^3); a quarter of each system's
# walkers use complex orbitals at a seeded twist (same cell and ions, own complex table).
UNEVEN = dict(nw=1024, sizes=(128, 256, 512, 1024), grids=(48, 64, 64, 80), complex_fraction=0.25, tau=0.3,
              sweeps=5, seed=20242)


def uneven_populations(nw: int = 1024, sizes=(128, 256, 512, 1024), grids=(48, 64, 64, 80),
                       complex_fraction: float = 0.25, seed: int = 20242, delay_rank: int = 32, **_ignored):
    """[(System, walkers, label)]: per size a real population and a complex one."""
    pops = []
    per = nw // len(sizes)
    for i, (N, M) in enumerate(zip(sizes, grids)):
        L = 18.0 * (N / 512.0) ** (1.0 / 3.0)
        kw = dict(L=L, n_ion=max(1, N // 4), n_up=N // 2, n_dn=N - N // 2, norb=N - N // 2, M=M,
                  delay_rank=min(delay_rank, N // 2), seed=seed + 10 * i)
        nc = int(round(per * complex_fraction))
        pops.append((make_system(**kw), per - nc, f"{N}e real"))
        if nc:
            pops.append((make_complex_system(**kw), nc, f"{N}e complex"))
    return pops


def make_walkers(system: System, nw: int, seed: int) -> np.ndarray:
    """Electron positions [nw][N][3], uniform in the cell, seeded."""
    rng = np.random.default_rng(seed + 7919)
    return rng.uniform(0.0, 1.0, size=(nw, system.N, 3)) * system.L[None, None, :]


def workload_system(name: str, **overrides) -> tuple[System, dict]:
    """The System of a named workload and its parameters (walkers, tau, sweeps, seed, ...)."""
    c = dict(WORKLOADS[name])
    c.update(overrides)
    return make_system(**c), c
