"""Orbitals from the reference's ESHDF test file on the HIP path, in the skewed (FCC) cell.

The diamond orbitals (tests/golden/eshdf/diamondC_1x1x1.pwscf.h5, converted by
qmcpack_amd/eshdf.py) are evaluated by the device (qmi_mw_spo_vgl, SPOSet::mw_evaluateVGL) at the
positions of QMCWaveFunctions/tests/test_einset_diamondC.cpp:59-124 after
qmi_system_set_lattice: against the reference's known answers (tolerances as in
tests/test_eshdf.py) and against the oracle's spline evaluation on the same table (1e-12).  With
the FP32 table the known answers hold at the reference's own single-precision level.  The
determinant path on the skewed cell: log|det| from the device's evaluateLog against numpy's
slogdet of Slater matrices filled by the oracle's spline.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import eshdf, qmi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "golden", "eshdf", "diamondC_1x1x1.pwscf.h5")


@pytest.fixture(scope="module")
def table():
    return eshdf.load_orbitals(H5, 8)


def context(table, precision="fp64", nw=2, seed=3):
    c = qmi.Context(0, precision=precision)
    c.system_create(eshdf.to_system(table, 4, 4, delay_rank=1))
    c.system_set_lattice(table.lattice)
    rng = np.random.default_rng(seed)
    R = rng.uniform(0, 1, size=(nw, 8, 3)) @ table.lattice
    c.walkers_load(R, 1)
    return c, R


def oracle_vgl(table, r):
    G = np.linalg.inv(table.lattice)
    u = np.asarray(r, float) @ G
    u -= np.floor(u)
    raw = orc.spline_vgh_raw(table.coefs, u)
    g = G @ raw[1:4]
    h = raw[4:10]
    H = np.array([[h[0], h[1], h[2]], [h[1], h[3], h[4]], [h[2], h[4], h[5]]])
    lap = np.einsum("ai,ijo,aj->o", G, H, G)
    return np.vstack([raw[0], g, lap])  # [5][norb]


KAT_POS = np.array([[0.0, 0.0, 0.0], [0.0, 1.0, 0.0]])


def check_kats(out, tol_v, tol_g, tol_l):
    v0, v1 = out[0, 0], out[1, 0]
    assert abs(v0[0] - -0.42546836868) < tol_v and abs(v0[1]) < tol_v
    assert abs(v1[0] - -0.8886948824) < tol_v and abs(v1[1] - 1.419412370359) < tol_v
    for q, want in enumerate([-0.0000183403, 0.1655139178, -0.0000193077]):
        assert abs(out[1, 1 + q, 0] - want) < tol_g
    for q, want in enumerate([-1.3131694794, -1.1174004078, -0.8462534547]):
        assert abs(out[1, 1 + q, 1] - want) < tol_g
    assert abs(out[1, 4, 0] - 1.3313053846) < tol_l * 1.3313053846
    assert abs(out[1, 4, 1] - -4.712583065) < tol_l * 4.712583065


def test_device_orbitals_match_known_answers(table):
    c, _ = context(table)
    out = c.mw_evaluateVGL(KAT_POS, 3)
    print("device diamond orbitals at r = (0,1,0):", out[1, :, :2].T)
    check_kats(out, 1e-7, 1e-6, 2e-5)
    for w in range(2):
        ref = oracle_vgl(table, KAT_POS[w])[:, :3]
        assert np.max(np.abs(out[w] - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # random positions anywhere in space (reduced and wrapped by the device)
    rng = np.random.default_rng(7)
    pos = rng.uniform(-12, 12, size=(2, 3))
    out = c.mw_evaluateVGL(pos, 8)
    for w in range(2):
        ref = oracle_vgl(table, pos[w])
        assert np.max(np.abs(out[w] - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))


def test_fp32_table_matches_known_answers(table):
    c, _ = context(table, precision="t32")
    out = c.mw_evaluateVGL(KAT_POS, 3)
    check_kats(out, 2e-6, 4e-6, 2e-5)


def test_determinant_on_skewed_cell(table):
    c, R = context(table, nw=3, seed=11)
    lp, G, L = c.mw_evaluateLog()
    for w in range(3):
        want = 0.0
        for g in range(2):
            rows = R[w, 4 * g:4 * g + 4]
            A = np.array([oracle_vgl(table, r)[0, :4] for r in rows])
            want += np.linalg.slogdet(A)[1]
        assert abs(lp[w] - want) < 1e-10 * max(1.0, abs(want))
    # a few determinant-only sweeps stay finite on the skewed cell
    E, _ = c.vmc_run(2, 1, 0.2, 0, 5)
    assert np.all(np.isfinite(E))


def test_skewed_cell_refusals(table):
    c, _ = context(table)
    with pytest.raises(qmi.QmiError):
        c.coulomb_setup(4.0)
    from qmcpack_amd.system import make_system
    s = make_system(L=6.0, n_ion=2, n_up=4, n_dn=4, norb=8, M=8, delay_rank=1, seed=1)
    c2 = qmi.Context(0)
    c2.system_create(s)
    with pytest.raises(qmi.QmiError):
        c2.system_set_lattice(table.lattice)  # Jastrow factors present
    c2.system_set_lattice(np.diag([6.0, 6.0, 6.0]))  # a diagonal cell is accepted
