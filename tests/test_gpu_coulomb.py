"""Periodic Coulomb energy on the HIP path against the CPU FP64 oracle, with both long-range
handlers: the Ewald breakup and the optimized LPQHI breakup (the reference's default).

The oracle's Ewald sums are pinned by tests/test_oracle_coulomb.py and its optimized breakup by
tests/test_oracle_breakup.py (the reference's known answers for CoulombPBCAA / CoulombPBCAB, the
Madelung constant, independence of the splitting).  Here the device's per-walker energies
(electron-electron, electron-ion, ion-ion) after a VMC run are compared with the oracle on the same
positions, within 1e-10 relative to max(|V|, 1).  For the optimized breakup the oracle evaluates
with the product's fitted coefficients (qmi_lr_breakup): the fit's near-null direction follows each
implementation's rounding (tests/test_breakup_host.py holds the fit itself to the oracle's).
"""
import copy

import numpy as np
import pytest

from oracle import breakup as bk
from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import make_complex_system, make_system, make_walkers

pytestmark = pytest.mark.gpu

REL = 1e-10


def potential_pair(system, R, zion, wpc=1, steps=1):
    c = qmi.Context(0)
    c.system_create(system)
    c.walkers_load(R, wpc)
    c.mw_evaluateLog()
    if steps:
        c.vmc_run(steps, 1, 0.4, 0, 17)
    c.coulomb_setup(zion)
    v = c.mw_evaluatePotential()
    o = orc.VMC(system, c.walkers_get(), wpc, 1)
    vo = o.potential(zion)
    return v, vo


def worst(v, vo):
    return float(np.max(np.abs(v - vo) / np.maximum(np.abs(vo), 1.0)))


@pytest.mark.parametrize("n_up,n_dn,nion,zion", [(5, 4, 3, 3.0), (20, 18, 8, 4.75), (64, 64, 16, 8.0)])
def test_coulomb_matches_oracle(n_up, n_dn, nion, zion):
    s = make_system(L=9.0 + 0.05 * nion, n_ion=nion, n_up=n_up, n_dn=n_dn, norb=max(n_up, n_dn), M=12,
                    delay_rank=4, seed=3 + nion)
    R = make_walkers(s, 5, 4)
    v, vo = potential_pair(s, R, zion, wpc=1 if nion != 8 else 2)
    print("coulomb", n_up + n_dn, "electrons, worst relative error", worst(v, vo), v[0], vo[0])
    assert worst(v, vo) < REL
    assert np.all(v[:, 2] == v[0, 2])


def test_coulomb_orthorhombic_cell_complex_orbitals():
    """An orthorhombic (unequal sides) cell, twisted complex orbitals: the energy depends only on
    positions, so the complex walkers give it the same way."""
    import dataclasses
    s = make_complex_system(L=10.0, n_ion=6, n_up=12, n_dn=10, norb=12, M=12, delay_rank=2, seed=9)
    s = dataclasses.replace(s, L=np.array([10.0, 11.5, 9.25]))
    R = make_walkers(s, 4, 5)
    v, vo = potential_pair(s, R, 2.0)
    assert worst(v, vo) < REL


def test_coulomb_production_size():
    """1024 electrons, 128 ions (the shape of the largest benchmark system): the row-paired
    electron-electron sum and the long-range sum over the k shells at full size."""
    s = make_system(L=22.0, n_ion=128, n_up=512, n_dn=512, norb=512, M=12, delay_rank=8, seed=2)
    R = make_walkers(s, 3, 6)
    v, vo = potential_pair(s, R, 4.0, steps=0)
    print("coulomb 1024 electrons, worst relative error", worst(v, vo))
    assert worst(v, vo) < REL


def test_coulomb_needs_setup_and_periodic_cell():
    s = make_system(L=9.0, n_ion=4, n_up=4, n_dn=4, norb=4, M=12, delay_rank=1, seed=1)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(make_walkers(s, 2, 1), 1)
    with pytest.raises(RuntimeError, match="qmi_coulomb_setup"):
        c.mw_evaluatePotential()


def oracle_opt_breakup(system, R, zion):
    """Per-walker {V_ee, V_eI, V_II} with the optimized breakup on the product's coefficients,
    short-range part through the 1001-point spline as on the device."""
    L = np.asarray(system.L, float) * np.ones(3)
    t, _ = qmi.lr_breakup(L)
    b = copy.copy(bk.Breakup(L))
    b.t = t
    ions = np.asarray(system.ions, float).reshape(-1, 3)
    qi = np.full(len(ions), zion)
    vii = bk.coulomb_aa(L, ions, qi, brk=b, spline=True)[0]
    out = []
    for r in R:
        qe = -np.ones(len(r))
        out.append([bk.coulomb_aa(L, r, qe, brk=b, spline=True)[0],
                    bk.coulomb_ab(L, ions, qi, r, qe, brk=b, spline=True)[0], vii])
    return np.array(out)


@pytest.mark.parametrize("n_up,n_dn,nion,zion,Lxyz", [(5, 4, 3, 3.0, None), (20, 18, 8, 4.75, (10.0, 11.5, 9.25)),
                                                      (64, 64, 16, 8.0, None)])
def test_opt_breakup_matches_oracle(n_up, n_dn, nion, zion, Lxyz):
    import dataclasses
    s = make_system(L=9.0 + 0.05 * nion, n_ion=nion, n_up=n_up, n_dn=n_dn, norb=max(n_up, n_dn), M=12,
                    delay_rank=4, seed=13 + nion)
    if Lxyz is not None:
        s = dataclasses.replace(s, L=np.array(Lxyz))
    R = make_walkers(s, 4, 14)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.mw_evaluateLog()
    c.vmc_run(1, 1, 0.4, 0, 19)
    c.coulomb_setup(zion, handler="opt_breakup")
    v = c.mw_evaluatePotential()
    Rw = c.walkers_get()
    vo = oracle_opt_breakup(s, Rw, zion)
    print("opt_breakup", n_up + n_dn, "electrons, worst relative error", worst(v, vo), v[0], vo[0])
    assert worst(v, vo) < REL
    # the two handlers approximate the same periodic energy: at the default cutoff (15) the fitted
    # split (through its spline) is within a few 1e-6 of the converged sum, the Ewald split only
    # ~1e-4 (real-space truncation erfc(sqrt(7.5))); against Ewald at cutoff 45 (converged to 1e-10)
    c.coulomb_setup(zion, dim_cutoff=45.0, handler="ewald")
    ve = c.mw_evaluatePotential()
    assert float(np.max(np.abs(v - ve) / np.maximum(np.abs(ve), 1.0))) < 1e-5


def test_opt_breakup_dmc_and_handler_validation():
    """DMC on the optimized breakup's tables (capacity re-allocation, branching), then the
    population's potential against the oracle; an unknown handler is rejected."""
    import ctypes as C
    s = make_system(L=7.0, n_ion=3, n_up=7, n_dn=6, norb=8, M=12, delay_rank=4, seed=65)
    R = make_walkers(s, 24, 66)
    zion = s.N / len(s.ions)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.coulomb_setup(zion, handler="opt_breakup")
    c.dmc_init(tau=0.02, capacity=64, seed=17, warmup_steps=2)
    st = c.dmc_run(4)
    assert np.all(np.isfinite(st))
    v = c.mw_evaluatePotential()
    vo = oracle_opt_breakup(s, c.walkers_get(), zion)
    assert worst(v, vo) < REL
    assert c.L.qmi_coulomb_setup_handler(c.h, C.c_double(1.0), C.c_double(15.0), C.c_int(7)) == 1


def test_structure_factor_matches_known_answer():
    """qmi_mw_structure_factor against StructFact's known answer (test_StructFact.cpp:24-70: cube
    of side 5, 3 + 1 particles, the 263786 k vectors up to k_c = 50, the sum over k of each
    species' rho_k within 5e-5) and against the direct sum at every k (1e-12)."""
    from qmcpack_amd.system import System
    K, _, _ = bk.lattice_kpoints([5.0, 5.0, 5.0], 50.0)
    R0 = np.array([[0.0, 1.0, 2.0], [1.0, 0.2, 3.0], [0.3, 4.0, 1.4], [3.2, 4.7, 0.7]])
    rng = np.random.default_rng(8)
    R = np.stack([R0, rng.uniform(0, 5, size=(4, 3))])
    s = System(L=np.array([5.0, 5.0, 5.0]), ions=np.zeros((0, 3)), n_up=3, n_dn=1, M=(0, 0, 0), norb=0, coefs=None,
               j1_params=None, j1_rcut=0.0, j1_cusp=0.0, j2_uu=None, j2_ud=None, j2_rcut=0.0, j2_cusp_uu=0.0,
               j2_cusp_ud=0.0, delay_rank=1)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    rho = c.mw_structure_factor(K)
    assert rho.shape == (2, 2, len(K))
    for g, want in enumerate([-125.80618630936, 68.199075127271]):
        assert abs(np.sum(rho[0, g]) - want) < 5e-5 * abs(want)
    for w in range(2):
        direct = [np.exp(1j * (K @ R[w, :3].T)).sum(1), np.exp(1j * (K @ R[w, 3:].T)).sum(1)]
        for g in range(2):
            assert np.max(np.abs(rho[w, g] - direct[g])) < 1e-12 * 3
