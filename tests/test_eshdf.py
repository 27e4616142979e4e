"""ESHDF orbital input (qmcpack_amd/h5read.py, qmcpack_amd/eshdf.py) against the reference's
known answers and the mathematics.

Fixture: tests/golden/eshdf/diamondC_1x1x1.pwscf.h5 is the reference's own test orbital file
(tests/solids/diamondC_1x1x1_pp/pwscf.pwscf.h5: two-atom diamond FCC primitive cell, Gamma point,
8 bands), data only.

Known answers: QMCWaveFunctions/tests/test_einset_diamondC.cpp:39-138 (the einspline builder on
that file, 8 orbitals, mesh factor 1): values, gradients, Laplacians and Hessians of the first
orbitals at r = (0, 0, 0) and (0, 1, 0).  The reference stores the table in single precision, so
its answers carry float rounding (its tiny gradient components 1.83e-5 / 1.93e-5 differ where the
crystal's symmetry makes them equal): values and gradients are held to 1e-7 relative to the
orbital's scale, the Laplacians and Hessian entries to the reference's own 1.2e-5 / 4.8e-4 (Catch
Approx, 100 and 2000 float epsilons).  The spline is evaluated by the oracle
(oracle.spline_vgh_raw, pinned elsewhere) at reduced coordinates u = r R^-1 wrapped into [0, 1),
gradients G g, Laplacians and Hessians through G = R^-1.

Mathematics: the spline interpolates the real-space orbital at every grid point; the phase
rotation leaves a real orbital; the spline stays within the interpolation error of the direct
plane-wave sum at random points; the mesh-size rule.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import eshdf, h5read

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "golden", "eshdf", "diamondC_1x1x1.pwscf.h5")
A = 3.37316115


@pytest.fixture(scope="module")
def table():
    return eshdf.load_orbitals(H5, 8)


def vgh(table, r):
    G = np.linalg.inv(table.lattice)
    u = np.asarray(r, float) @ G
    u -= np.floor(u)
    raw = orc.spline_vgh_raw(table.coefs, u)
    v, g, h = raw[0], raw[1:4], raw[4:10]
    grad = G @ g
    H = np.array([[h[0], h[1], h[2]], [h[1], h[3], h[4]], [h[2], h[4], h[5]]])  # [3][3][norb]
    hess = np.einsum("ai,ijo,bj->abo", G, H, G)
    lap = np.einsum("aao->o", hess)
    return v, grad, lap, hess


def test_reader(table):
    f = h5read.File(H5)
    assert f.keys("/") == ["application", "atoms", "electrons", "format", "supercell", "version"]
    assert np.array_equal(f["/version"], [2, 0, 0]) and f["/format"][0] == b"ES-HDF"
    assert np.allclose(f["/supercell/primitive_vectors"], [[A, A, 0], [0, A, A], [A, 0, A]], atol=0, rtol=0)
    assert np.allclose(f["/atoms/positions"], [[0, 0, 0], [1.68658058] * 3], rtol=1e-9)
    assert f["/electrons/kpoint_0/gvectors"].shape == (3695, 3)
    assert f["/electrons/kpoint_0/spin_0/state_7/psi_g"].shape == (3695, 2)
    assert int(f["/electrons/number_of_kpoints"][0]) == 1
    assert table.mesh == (40, 40, 40) and list(table.bands) == list(range(8))
    assert np.all(np.diff(table.energies) >= 0)
    with pytest.raises(h5read.H5Error):
        h5read.File(os.path.join(HERE, "golden", "reference_kats.json"))


def test_known_answers_diamond(table):
    v0, g0, l0, _ = vgh(table, [0.0, 0.0, 0.0])
    v1, g1, l1, h1 = vgh(table, [0.0, 1.0, 0.0])
    close = lambda a, b, tol: abs(a - b) <= tol
    assert close(v0[0], -0.42546836868, 1e-7) and close(v0[1], 0.0, 1e-7)
    assert close(v1[0], -0.8886948824, 1e-7) and close(v1[1], 1.419412370359, 1e-7)
    for got, want in zip(g1[:, 0], [-0.0000183403, 0.1655139178, -0.0000193077]):
        assert close(got, want, 1e-6)
    for got, want in zip(g1[:, 1], [-1.3131694794, -1.1174004078, -0.8462534547]):
        assert close(got, want, 1e-6)
    approx = 100 * np.finfo(np.float32).eps
    assert close(l1[0], 1.3313053846, 2e-5 * 1.3313053846)
    assert close(l1[1], -4.712583065, approx * 4.712583065)
    want = [[-2.3160984034, 1.8089479397, 0.5608575749], [1.8089479397, -0.07996207476, 0.5237969314],
            [0.5608575749, 0.5237969314, -2.316497764]]
    eps = 2000 * np.finfo(np.float32).eps
    for a in range(3):
        for b in range(3):
            tol = eps * 0.08 if (a, b) == (1, 1) else approx * abs(want[a][b])
            assert close(h1[a, b, 1], want[a][b], max(tol, 2e-6)), (a, b)


def test_interpolation_and_plane_wave_sum(table):
    f = h5read.File(H5)
    gv = f["/electrons/kpoint_0/gvectors"].astype(np.int64)
    # the spline passes through the real-space orbital at the grid points
    box = eshdf.real_space_band(f["/electrons/kpoint_0/spin_0/state_2/psi_g"], gv, table.mesh)
    real = eshdf.rotate_to_real(box, [0, 0, 0])
    rng = np.random.default_rng(5)
    M = np.array(table.mesh)
    for _ in range(6):
        idx = rng.integers(0, M)
        raw = orc.spline_vgh_raw(table.coefs, idx / M)
        assert abs(raw[0][2] - real[tuple(idx)]) < 1e-12 * np.max(np.abs(real))
    # the lowest band is non-degenerate: a real function times a global phase, which the rotation
    # removes (a degenerate band can be a complex combination; the rotation then keeps the real
    # part of its best phase, as the reference does)
    psi = f["/electrons/kpoint_0/spin_0/state_0/psi_g"]
    box0 = eshdf.real_space_band(psi, gv, table.mesh)
    z = box0 / np.exp(1j * np.angle(np.sum(box0 * box0)) / 2)
    assert np.max(np.abs(z.imag)) < 1e-10 * np.max(np.abs(z.real))
    real0 = eshdf.rotate_to_real(box0, [0, 0, 0])
    # the output is one unit-modulus rotation of z: a fixed real multiple (|m| <= 1) of the orbital
    big = np.abs(z.real) > 0.1 * np.max(np.abs(z.real))
    m = real0[big] / z.real[big]
    assert np.ptp(m) < 1e-10 and 0 < abs(m[0]) <= 1.0
    # against the direct plane-wave sum at random points (the mesh's interpolation error)
    c = psi[:, 0] + 1j * psi[:, 1]
    for _ in range(4):
        u = rng.uniform(0, 1, 3)
        direct = abs(np.sum(c * np.exp(2j * np.pi * (gv @ u))))
        raw = orc.spline_vgh_raw(table.coefs, u)
        assert abs(abs(raw[0][0]) - abs(m[0]) * direct) < 2e-3 * np.max(np.abs(real0))


def test_mesh_size_rule():
    g = lambda m: np.array([[m, 0, 0], [0, m, 0], [0, 0, m]])
    assert eshdf.mesh_size(g(10)) == (40, 40, 40)
    assert eshdf.mesh_size(g(11)) == (44, 44, 44)
    assert eshdf.mesh_size(g(10), 0.55) == (22, 22, 22)
    assert eshdf.mesh_size(g(33)) == (135, 135, 135)  # 132 -> next 2^a 3^b 5^c
    assert eshdf.mesh_size(g(40)) == (160, 160, 160)
    assert len(eshdf._ALLOWED) == 257 and eshdf._ALLOWED[0] == 72 and eshdf._ALLOWED[-1] == 65536
