"""Orbital input from ESHDF files (the plane-wave converters' HDF5 orbital format) into periodic
B-spline tables for the device.

Follows the reference's einspline orbital builder for one spin and one twist of the primitive cell
(tilematrix = identity):
  * cell, atoms, bands: EinsplineSetBuilderESHDF.fft.cpp:71-110, 250-262, 490-505 (lattice rows
    "/supercell/primitive_vectors", "/atoms/positions", per-twist "reduced_k", "gvectors",
    "spin_s/eigenvalues", "state_i/psi_g");
  * band order: ascending energy, ties (within 1e-6) by band index (BandInfo.h:50-60);
  * mesh: "/electrons/psi_r_mesh" or "/electrons/mesh" if present, else ceil(4 * meshfactor *
    max |G_i|) per direction, rounded up to even at or below 128 and to the next 5-smooth size
    (2^a 3^b 5^c) above (EinsplineSetBuilderESHDF.fft.cpp:366-436);
  * per band: the coefficients placed on the FFT box at G mod mesh for G inside the box
    (einspline_helper.hpp:34-59 unpack4fftw), an unnormalised inverse FFT (FFTW sign +1), the twist
    phase exp(-2 pi i u.k) applied and the global phase rotated so that the result is real
    (einspline_helper.hpp:70-127 fix_phase_rotate_c2r: a unit-modulus rotation chosen from the
    real / imaginary norms; an orbital that is already real comes out scaled by 1/sqrt(2), which
    the reference's known answers carry and determinant ratios do not see), and the periodic cubic
    B-spline
    interpolating those grid values on [0, 1)^3 (einspline's periodic find_coefs: at every grid
    point c_{i-1} / 6 + 2 c_i / 3 + c_{i+1} / 6 equals the value), solved here per direction in
    Fourier space (the system is circulant).

Returns the unique periodic coefficient table [Mx][My][Mz][norb] the system descriptor takes, the
lattice rows R (r = u R), the atoms and the per-band energies.  Host-side setup, numpy only.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .h5read import File


def _five_smooth(lo=72, hi=65536):
    out = []
    for x in range(lo, hi + 1):
        y = x
        for p in (2, 3, 5):
            while y % p == 0:
                y //= p
        if y == 1:
            out.append(x)
    return np.array(out)


_ALLOWED = _five_smooth()


def mesh_size(gvecs: np.ndarray, meshfactor: float = 1.0) -> tuple[int, int, int]:
    maxg = np.max(np.abs(gvecs), axis=0)
    out = []
    for m in maxg:
        n = int(np.ceil(4.0 * meshfactor * m))
        out.append(int(_ALLOWED[np.searchsorted(_ALLOWED, n)]) if n > 128 else n + n % 2)
    return tuple(out)


def real_space_band(psi_g: np.ndarray, gvecs: np.ndarray, mesh) -> np.ndarray:
    """Unnormalised sum_G c_G exp(+2 pi i G.u) on the mesh (u = index / mesh)."""
    mesh = np.asarray(mesh)
    upper = (mesh - 1) // 2
    lower = upper - mesh + 1
    inside = np.all((gvecs <= upper) & (gvecs >= lower), axis=1)
    box = np.zeros(tuple(mesh), dtype=np.complex128)
    idx = np.mod(gvecs[inside], mesh)
    box[idx[:, 0], idx[:, 1], idx[:, 2]] = psi_g[inside, 0] + 1j * psi_g[inside, 1]
    return np.fft.ifftn(box) * box.size


def rotate_to_real(box: np.ndarray, twist) -> np.ndarray:
    """Multiply by exp(-2 pi i u.k) and rotate the global phase so that the orbital is real; returns
    phase_r Re - phase_i Im with the phase from the real / imaginary norms."""
    n = box.shape
    u = [np.arange(n[d]) / n[d] * twist[d] for d in range(3)]
    arg = -2.0 * np.pi * (u[0][:, None, None] + u[1][None, :, None] + u[2][None, None, :])
    z = box * np.exp(1j * arg)
    rn = float(np.sum(z.real * z.real))
    inn = float(np.sum(z.imag * z.imag))
    rin = float(np.sum(z.real * z.imag))
    x = (rn - inn) / rin
    y = 1.0 / np.sqrt(x * x + 4.0)
    pr = np.sqrt(0.5 - y)
    pi = np.sqrt(1.0 - pr * pr) * (1.0 if x < 0 else -1.0)
    return pr * z.real - pi * z.imag


def periodic_spline_coefs(values: np.ndarray) -> np.ndarray:
    """Coefficients c of the periodic cubic B-spline through values on a uniform periodic grid
    (the first three axes): (c_{i-1} + 4 c_i + c_{i+1}) / 6 = v_i along each axis."""
    c = np.asarray(values, dtype=np.float64)
    for ax in range(3):
        n = c.shape[ax]
        lam = (4.0 + 2.0 * np.cos(2.0 * np.pi * np.arange(n) / n)) / 6.0
        shape = [1] * c.ndim
        shape[ax] = n
        c = np.fft.ifft(np.fft.fft(c, axis=ax) / lam.reshape(shape), axis=ax).real
    return np.ascontiguousarray(c)


@dataclasses.dataclass
class OrbitalTable:
    coefs: np.ndarray        # [Mx][My][Mz][norb] unique periodic coefficients
    lattice: np.ndarray      # rows are the cell vectors (bohr)
    ions: np.ndarray         # [n_ion][3] Cartesian (bohr)
    species: np.ndarray      # [n_ion] species index
    energies: np.ndarray     # [norb] band energies, ascending
    bands: np.ndarray        # [norb] band index in the file
    twist: np.ndarray        # reduced k of the twist
    mesh: tuple


def load_orbitals(path: str, norb: int, spin: int = 0, twist: int = 0, meshfactor: float = 1.0) -> OrbitalTable:
    """The norb lowest bands of one spin at one twist as a periodic B-spline table."""
    f = File(path)
    lattice = f["/supercell/primitive_vectors"].astype(np.float64)
    nk = int(np.ravel(f["/electrons/number_of_kpoints"])[0])
    if not 0 <= twist < nk:
        raise ValueError(f"twist {twist} not in the file ({nk} twists)")
    kp = f"/electrons/kpoint_{twist}"
    gvecs = f[f"{kp}/gvectors"].astype(np.int64)
    g0 = f["/electrons/kpoint_0/gvectors"].astype(np.int64)
    mesh = None
    for name in ("/electrons/psi_r_mesh", "/electrons/mesh"):
        if name in f:
            mesh = tuple(int(x) for x in np.ravel(f[name]))
    if mesh is None:
        mesh = mesh_size(g0, meshfactor)
    eig = np.ravel(f[f"{kp}/spin_{spin}/eigenvalues"]).astype(np.float64)
    nstates = int(np.ravel(f[f"{kp}/spin_{spin}/number_of_states"])[0])
    if norb > nstates:
        raise ValueError(f"{norb} orbitals requested, {nstates} in the file")
    # ascending energy; runs of energies within 1e-6 of each other in band order
    order = sorted(range(nstates), key=lambda b: eig[b])
    i = 0
    while i < nstates:
        j = i + 1
        while j < nstates and eig[order[j]] - eig[order[j - 1]] < 1e-6:
            j += 1
        order[i:j] = sorted(order[i:j])
        i = j
    bands = np.array(order[:norb])
    kred = np.ravel(f[f"{kp}/reduced_k"]).astype(np.float64)
    table = np.empty(tuple(mesh) + (norb,))
    for o, b in enumerate(bands):
        psi_g = f[f"{kp}/spin_{spin}/state_{b}/psi_g"].astype(np.float64)
        box = real_space_band(psi_g, gvecs, mesh)
        table[..., o] = rotate_to_real(box, kred)
    coefs = periodic_spline_coefs(table)
    ions = f["/atoms/positions"].astype(np.float64)
    species = np.ravel(f["/atoms/species_ids"]).astype(np.int64)
    return OrbitalTable(coefs=coefs, lattice=lattice, ions=ions, species=species, energies=eig[bands], bands=bands,
                        twist=kred, mesh=tuple(mesh))


def orthorhombic(lattice: np.ndarray) -> bool:
    return bool(np.all(lattice[~np.eye(3, dtype=bool)] == 0.0))


def to_system(table: OrbitalTable, n_up: int, n_dn: int, delay_rank: int = 1):
    """A system description for the table's orbitals without Jastrow factors.  For a cell with
    off-diagonal vectors the caller sets the lattice after system_create
    (Context.system_set_lattice(table.lattice)); the description's sides are then the vector
    lengths, used by nothing that accepts such a cell."""
    from .system import System
    L = np.diag(table.lattice).copy() if orthorhombic(table.lattice) else np.linalg.norm(table.lattice, axis=1)
    return System(L=L, ions=table.ions.copy(), n_up=n_up, n_dn=n_dn, M=tuple(table.mesh), norb=table.coefs.shape[3],
                  coefs=table.coefs, j1_params=None, j1_rcut=0.0, j1_cusp=0.0, j2_uu=None, j2_ud=None, j2_rcut=0.0,
                  j2_cusp_uu=0.0, j2_cusp_ud=0.0, delay_rank=delay_rank)
