"""Non-local pseudopotential energy on the HIP path against the CPU FP64 oracle.

The oracle's energy is pinned by tests/test_oracle_nlpp.py (constant psi gives the l = 0 channel
sum, spherical psi around one ion, the angular assembly in numpy with scipy's Legendre
polynomials).  Here the device evaluates the same walkers after VMC runs (completed delayed
updates) with the same radial splines, quadrature rule and per-walker rotations: 1e-10 relative
to max(|V|, 1).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import make_complex_system, make_system, make_walkers

from test_oracle_nlpp import bare_system, channels, icosahedron, random_rotations

pytestmark = pytest.mark.gpu

REL = 1e-10


def nlpp_pair(s, nw, ls, rcut, steps=1, seed=31, wpc=1):
    R = make_walkers(s, nw, 3)
    dr, v, v2, _ = channels(ls, seed=len(ls))
    knots, w = icosahedron()
    rot = random_rotations(nw, seed)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, wpc)
    c.mw_evaluateLog()
    if steps:
        c.vmc_run(steps, 1, 0.4, 0, seed)
    c.nlpp_setup(ls, dr, v, v2, rcut, knots, w)
    E = c.mw_evaluateNLPP(rot)
    o = orc.VMC(s, c.walkers_get(), 1, 1)
    o.evaluate_log()
    Eo = o.nlpp_energy(ls, dr, v, v2, rcut, knots, w, rot)
    return c, E, Eo


def worst(E, Eo):
    return float(np.max(np.abs(E - Eo) / np.maximum(np.abs(Eo), 1.0)))


@pytest.mark.parametrize("k,ls", [(1, [0, 1, 2]), (4, [0, 1]), (1, [0, 1, 2, 3])])
def test_nlpp_matches_oracle(k, ls):
    s = make_system(L=7.0, n_ion=6, n_up=8, n_dn=7, norb=9, M=12, delay_rank=k, seed=11)
    c, E, Eo = nlpp_pair(s, 6, ls, 2.2, steps=2)
    print("nlpp", E, Eo, worst(E, Eo))
    assert np.any(np.abs(Eo) > 1e-3)
    assert worst(E, Eo) < REL


def test_nlpp_constant_psi():
    rng = np.random.default_rng(1)
    L = np.array([6.0, 6.5, 7.0])
    s = bare_system(6, 5, rng.uniform(0, 1, size=(5, 3)) * L, L)
    c, E, Eo = nlpp_pair(s, 4, [0, 1, 2, 3], 2.5, steps=0)
    assert worst(E, Eo) < REL


@pytest.mark.parametrize("n_up,n_dn", [(128, 128), (300, 270)])
def test_nlpp_large_determinants(n_up, n_dn):
    """256 orbitals (one per thread) and 300 (two per thread: the inverse row in registers)."""
    s = make_system(L=14.0, n_ion=16, n_up=n_up, n_dn=n_dn, norb=max(n_up, n_dn), M=16, delay_rank=16, seed=5)
    c, E, Eo = nlpp_pair(s, 2, [0, 1], 1.6, steps=0)
    print("nlpp large", E, Eo, worst(E, Eo))
    assert worst(E, Eo) < 1e-9


def test_nlpp_refusals():
    s = make_complex_system(L=7.0, n_ion=3, n_up=4, n_dn=4, norb=4, M=12, delay_rank=1, seed=2)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(make_walkers(s, 2, 1), 1)
    with pytest.raises(RuntimeError, match="qmi_nlpp_setup"):
        c.mw_evaluateNLPP(random_rotations(2, 1))
    dr, v, v2, _ = channels([0])
    knots, w = icosahedron()
    with pytest.raises(RuntimeError, match="r_cut inside the grid"):
        c.nlpp_setup([0], dr, v, v2, 100.0, knots, w)
    c.nlpp_setup([0], dr, v, v2, 2.0, knots, w)
    with pytest.raises(RuntimeError, match="real orbitals"):
        c.mw_evaluateNLPP(random_rotations(2, 1))


@pytest.mark.parametrize("prec", ["fp32", "t32"])
def test_nlpp_mixed_precision(prec):
    """The FP32-table and FP32-engine forms against the oracle on the same positions: the table
    (and, for fp32, the inverse) rounded to FP32 moves the energy by ~1e-7 relative (measured
    1.1e-7 and 1.0e-7)."""
    s = make_system(L=7.0, n_ion=6, n_up=8, n_dn=7, norb=9, M=12, delay_rank=4, seed=11)
    R = make_walkers(s, 4, 3)
    dr, v, v2, _ = channels([0, 1], seed=2)
    knots, w = icosahedron()
    rot = random_rotations(4, 9)
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.mw_evaluateLog()
    c.nlpp_setup([0, 1], dr, v, v2, 2.2, knots, w)
    E = c.mw_evaluateNLPP(rot)
    o = orc.VMC(s, R, 1, 1)
    o.evaluate_log()
    Eo = o.nlpp_energy([0, 1], dr, v, v2, 2.2, knots, w, rot)
    print(prec, worst(E, Eo))
    assert worst(E, Eo) < 1e-4


def energies_by_form(monkeypatch, s, R, prec, ls, rcut, seed, form):
    """mw_evaluateNLPP on fixed positions with the evaluation form chosen at set-up."""
    if form == "electron":
        monkeypatch.setenv("QMI_NLPP_FORM", "electron")
    else:
        monkeypatch.delenv("QMI_NLPP_FORM", raising=False)
    dr, v, v2, _ = channels(ls, seed=2)
    knots, w = icosahedron()
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.mw_evaluateLog()
    c.nlpp_setup(ls, dr, v, v2, rcut, knots, w)
    rot = random_rotations(R.shape[0], seed)
    E = c.mw_evaluateNLPP(rot)
    E2 = c.mw_evaluateNLPP(rot)  # the pair list is rebuilt each call
    assert np.array_equal(E, E2)
    return E, (dr, v, v2, knots, w, rot)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pair_list_form_matches_per_electron_form(monkeypatch, prec):
    """The pair-list form (the electron-ion pairs of all walkers sorted by ion and evaluated by a
    persistent grid, the pairs of one ion together on one XCD) against one workgroup per electron:
    the same per-pair code and the same summation order (points, then ions in ion order per
    electron, then electrons); the two kernels differ only where the compiler fuses multiply-adds
    (measured 2e-15 relative), so within 1e-13 relative to max(|V|, 1).  Shaped like the benchmark
    system: 512 electrons at its density, 128 ions, s and p channels, r_cut 1.5, 12 points."""
    s = make_system(L=18.0, n_ion=128, n_up=256, n_dn=256, norb=256, M=16, delay_rank=32, seed=31)
    R = make_walkers(s, 24, 5)
    Ep, _ = energies_by_form(monkeypatch, s, R, prec, [0, 1], 1.5, 3, "pairs")
    Ee, _ = energies_by_form(monkeypatch, s, R, prec, [0, 1], 1.5, 3, "electron")
    print(prec, "pair-list vs per-electron:", float(np.max(np.abs(Ep - Ee))), "energies", Ep[:3])
    assert np.all(Ep != 0.0)
    assert worst(Ep, Ee) < 1e-13


def test_pair_list_with_clustered_ions(monkeypatch):
    """Seven ions within 0.5 bohr (per coordinate) of one point and five electrons of every walker
    placed among them, each within r_cut of all seven: the pair-list form against the
    per-electron form (1e-13) and the oracle (1e-10)."""
    rng = np.random.default_rng(8)
    s = make_system(L=8.0, n_ion=10, n_up=9, n_dn=8, norb=9, M=12, delay_rank=4, seed=17)
    ions = s.ions.copy()
    ions[:7] = 4.0 + rng.uniform(-0.5, 0.5, size=(7, 3))
    import dataclasses
    s = dataclasses.replace(s, ions=ions)
    R = make_walkers(s, 6, 9)
    R[:, :5] = 4.0 + rng.uniform(-0.6, 0.6, size=(6, 5, 3))
    Ep, (dr, v, v2, knots, w, rot) = energies_by_form(monkeypatch, s, R, "fp64", [0, 1, 2], 2.0, 4, "pairs")
    Ee, _ = energies_by_form(monkeypatch, s, R, "fp64", [0, 1, 2], 2.0, 4, "electron")
    o = orc.VMC(s, R, 1, 1)
    o.evaluate_log()
    Eo = o.nlpp_energy([0, 1, 2], dr, v, v2, 2.0, knots, w, rot)
    print("clustered ions:", Ep, Eo, worst(Ep, Eo))
    assert worst(Ep, Ee) < 1e-13
    assert worst(Ep, Eo) < REL


@pytest.mark.parametrize("form", ["pairs", "electron"])
def test_nlpp_without_ions(monkeypatch, form):
    """A system with no ions has no electron-ion pairs: zero energy in both forms (the pair list's
    per-ion arrays are empty)."""
    from test_oracle_nlpp import bare_system
    L = np.array([6.0, 6.5, 7.0])
    s = bare_system(4, 3, np.zeros((0, 3)), L)
    R = np.random.default_rng(2).uniform(0, 1, size=(3, s.N, 3)) * L
    E, _ = energies_by_form(monkeypatch, s, R, "fp64", [0, 1], 2.0, 5, form)
    assert np.array_equal(E, np.zeros(3))
