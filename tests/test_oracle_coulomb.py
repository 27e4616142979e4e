"""Pin the oracle's periodic Coulomb energy (Ewald breakup, EwaldHandler3D with CoulombPBCAA /
CoulombPBCAB) against the reference's own known answers and the mathematics.

Known answers: QMCHamiltonians/tests/test_coulomb_pbcAA_ewald.cpp:28-135 (one charge in a unit
cube, two in a BCC cell) and test_coulomb_pbcAB_ewald.cpp:28-160 (A-B constants, energies and the
total), quoted there to 6-7 digits (checked here at 1e-6 relative).  Mathematics: the Madelung
energy of the simple cubic lattice (-2.8372974795 / 2 per unit charge) approached as the k-space
cutoff grows; translation invariance; the A-B decomposition of a two-species system equals the
A-A energy of the union (constants included); independence of the splitting parameter (the total
at two different cutoffs agrees while its short- and long-range parts move).
"""
import numpy as np

from oracle import oracle as orc

SC = np.array([1.0, 1.0, 1.0])
A_BCC = 3.77945227
BCC = np.array([A_BCC] * 3)
BCC_IONS = np.array([[0.0, 0.0, 0.0], [1.88972614, 1.88972614, 1.88972614]])


def close(a, b, rel=1e-6):
    return abs(a - b) <= rel * abs(b)


def test_ewald_cutoffs():
    rc, kc, sigma, nk = orc.ewald_params(SC)
    assert rc == 0.5 and abs(kc - 30.0) < 1e-12 and abs(sigma - np.sqrt(30.0)) < 1e-12 and nk > 0


def test_aa_known_answers():
    v, parts = orc.coulomb_aa(SC, [[0.0, 0.0, 0.0]], [1.0])
    assert close(parts[2], -3.142553) and close(v, -1.418927)
    v, parts = orc.coulomb_aa(SC, [[0.0, 0.5, 0.0]], [-1.0])
    assert close(parts[2], -3.142553) and close(v, -1.418927)
    v, parts = orc.coulomb_aa(BCC, BCC_IONS, [1.0, 1.0])
    assert close(parts[2], -1.690675) and close(v, -0.963074)


def test_ab_known_answers():
    vs = orc.coulomb_ab(SC, [[0.0, 0.0, 0.0]], [1.0], [[0.5, 0.0, 0.0]], [-1.0])
    assert close(vs[1][2], 0.0523598776 * 2) and close(vs[0], -0.008302 + 0.0523598776 * 2, 1e-5)
    ee = orc.coulomb_aa(SC, [[0.5, 0.0, 0.0]], [-1.0])[0]
    ii = orc.coulomb_aa(SC, [[0.0, 0.0, 0.0]], [1.0])[0]
    assert close(ee, -1.418927) and close(ii, -1.418927) and close(ee + ii + vs[0], -2.741436)
    elec = np.array([[0.5, 0.0, 0.0], [0.0, 0.5, 0.0]])
    vei, p = orc.coulomb_ab(BCC, BCC_IONS, [1.0, 1.0], elec, [-1.0, -1.0])
    assert close(p[2], 0.0277076538 * 4) and close(vei, -2.223413 + 0.0277076538 * 4)
    vee = orc.coulomb_aa(BCC, elec, [-1.0, -1.0])[0]
    vii = orc.coulomb_aa(BCC, BCC_IONS, [1.0, 1.0])[0]
    assert close(vee, -0.012808 - 0.0277076538 * 2, 1e-5) and close(vii, -0.907659 - 0.0277076538 * 2)
    assert close(vee + vii + vei, -3.143880)


def test_madelung_limit_and_splitting_independence():
    exact = -2.837297479480620 / 2
    errs = [abs(orc.coulomb_aa(SC, [[0.1, 0.2, 0.3]], [1.0], dim_cutoff=c)[0] - exact) for c in (15.0, 25.0, 40.0)]
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-7
    # a random neutral cell: the total barely moves with the cutoff while its parts do
    rng = np.random.default_rng(3)
    L = np.array([5.0, 6.0, 7.0])
    r = rng.uniform(0, 1, size=(10, 3)) * L
    q = np.array([1.0, -1.0] * 5)
    # (at these cutoffs both the real-space truncation erfc(sigma r_c) and exp(-k_c^2 / 4 sigma^2)
    # are below 1e-10)
    v1, p1 = orc.coulomb_aa(L, r, q, dim_cutoff=45.0)
    v2, p2 = orc.coulomb_aa(L, r, q, dim_cutoff=60.0)
    assert abs(v1 - v2) < 1e-8 * max(1.0, abs(v1)) and abs(p1[1] - p2[1]) > 0.1


def test_translation_invariance_and_ab_decomposition():
    rng = np.random.default_rng(4)
    L = np.array([6.0, 7.0, 8.0])
    re = rng.uniform(0, 1, size=(7, 3)) * L
    ri = rng.uniform(0, 1, size=(3, 3)) * L
    qe, qi = -np.ones(7), np.array([2.0, 3.0, 2.0])
    vee, vii = orc.coulomb_aa(L, re, qe)[0], orc.coulomb_aa(L, ri, qi)[0]
    vei = orc.coulomb_ab(L, ri, qi, re, qe)[0]
    vall = orc.coulomb_aa(L, np.vstack([re, ri]), np.concatenate([qe, qi]))[0]
    assert abs(vee + vii + vei - vall) < 1e-10 * abs(vall)
    t = np.array([0.37, -1.2, 2.9])
    assert abs(orc.coulomb_aa(L, re + t, qe)[0] - vee) < 1e-10 * abs(vee)
    assert abs(orc.coulomb_ab(L, ri + t, qi, re + t, qe)[0] - vei) < 1e-10 * abs(vei)
