"""The oracle's non-local pseudopotential energy (NonLocalECPotential::evaluate with
NonLocalECPComponent::evaluateOne) pinned independently of its own arithmetic:

* psi = 1: every ratio is 1 and the 12-point icosahedral rule integrates P_l exactly for l <= 5
  (sum_j w_j P_l(u_j . e) = delta_l0), so V = sum over electron-ion pairs within the cutoff of
  v_0(r) -- with v_0 a natural cubic spline built by scipy, the pairs found by numpy;
* one ion and a one-body Jastrow only: psi is constant on every sphere around the ion, the same
  sum holds;
* a full wave function (determinants, J1, J2): the angular assembly recomputed in numpy from the
  oracle's single-move ratios (pinned against recomputation of log psi in tests/test_oracle_math.py)
  and scipy's Legendre polynomials.
"""
import numpy as np
from scipy.interpolate import CubicSpline
from scipy.special import eval_legendre

from oracle import oracle as orc
from qmcpack_amd.system import System, jastrow_params, make_system, make_walkers


def icosahedron():
    g = (1 + 5 ** 0.5) / 2
    q = np.array([[0, 1, g], [0, -1, g], [0, 1, -g], [0, -1, -g], [1, g, 0], [-1, g, 0], [1, -g, 0], [-1, -g, 0],
                  [g, 0, 1], [-g, 0, 1], [g, 0, -1], [-g, 0, -1]], dtype=float)
    return q / np.linalg.norm(q, axis=1)[:, None], np.full(12, 1.0 / 12)


def channels(ls, rmax=3.0, ng=121, seed=0):
    """Radial channels as natural cubic splines on r_i = i dr: values and second derivatives."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, rmax, ng)
    splines, v, v2 = [], [], []
    for i, _ in enumerate(ls):
        a, b = rng.uniform(0.5, 2.0), rng.uniform(0.5, 1.5)
        y = a * np.exp(-b * x * x) * (1 + 0.3 * np.cos(2 * x + i))
        cs = CubicSpline(x, y, bc_type="natural")
        splines.append(cs)
        v.append(y)
        v2.append(cs(x, 2))
    return x[1] - x[0], np.array(v), np.array(v2), splines


def random_rotations(nw, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nw):
        A, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        out.append(A)
    return np.array(out)


def pairs(L, ions, R, rcut):
    """(electron, ion, r, r_e - r_I) within the cutoff, minimum image, electron-major order."""
    out = []
    for e, re in enumerate(R):
        for a, ri in enumerate(ions):
            d = re - ri
            d -= L * np.round(d / L)
            r = np.linalg.norm(d)
            if r < rcut:
                out.append((e, a, r, d))
    return out


def bare_system(n_up, n_dn, ions, L, with_j1=False):
    p = jastrow_params(8)
    return System(L=L, ions=ions, n_up=n_up, n_dn=n_dn, M=(0, 0, 0), norb=0, coefs=None,
                  j1_params=p if with_j1 else None, j1_rcut=2.5, j1_cusp=0.0, j2_uu=None, j2_ud=None, j2_rcut=0.0,
                  j2_cusp_uu=0.0, j2_cusp_ud=0.0, delay_rank=1)


def test_constant_psi_gives_l0_channel_sum():
    rng = np.random.default_rng(1)
    L = np.array([6.0, 6.5, 7.0])
    ions = rng.uniform(0, 1, size=(5, 3)) * L
    s = bare_system(6, 5, ions, L)
    R = rng.uniform(0, 1, size=(4, s.N, 3)) * L
    ls = [0, 1, 2, 3]
    dr, v, v2, sp = channels(ls)
    knots, w = icosahedron()
    rot = random_rotations(4, 2)
    o = orc.VMC(s, R, 1, 1)
    E = o.nlpp_energy(ls, dr, v, v2, 2.5, knots, w, rot)
    for iw in range(4):
        P = pairs(L, ions, R[iw], 2.5)
        assert len(P) > 2
        ref = sum(float(sp[0](r)) for _, _, r, _ in P)
        assert abs(E[iw] - ref) < 1e-12 * max(1.0, abs(ref)), (E[iw], ref)


def test_one_ion_j1_only_is_spherical():
    rng = np.random.default_rng(2)
    L = np.array([7.0, 7.0, 7.0])
    ions = np.array([[3.5, 3.4, 3.3]])
    s = bare_system(5, 4, ions, L, with_j1=True)
    R = ions[0] + rng.normal(size=(3, s.N, 3)) * 1.2
    ls = [0, 2]
    dr, v, v2, sp = channels(ls, seed=3)
    knots, w = icosahedron()
    o = orc.VMC(s, R, 1, 1)
    o.evaluate_log()
    E = o.nlpp_energy(ls, dr, v, v2, 2.4, knots, w, random_rotations(3, 4))
    for iw in range(3):
        ref = sum(float(sp[0](r)) for _, _, r, _ in pairs(L, ions, R[iw], 2.4))
        assert abs(E[iw] - ref) < 1e-11 * max(1.0, abs(ref))


def test_full_wave_function_against_numpy_assembly():
    s = make_system(L=6.0, n_ion=3, n_up=5, n_dn=4, norb=6, M=12, delay_rank=1, seed=7)
    R = make_walkers(s, 2, 8)
    ls = [0, 1, 2]
    dr, v, v2, sp = channels(ls, seed=5)
    knots, w = icosahedron()
    rot = random_rotations(2, 6)
    o = orc.VMC(s, R, 1, 1)
    o.evaluate_log()
    E = o.nlpp_energy(ls, dr, v, v2, 2.0, knots, w, rot)
    for iw in range(2):
        u = knots @ rot[iw].T
        tot = 0.0
        P = pairs(s.L, s.ions, R[iw], 2.0)
        assert len(P) >= 3
        for e, a, r, d in P:
            ratio = np.array([o.ratio_grad(iw, e, R[iw, e] + (r * uj - d))[0] for uj in u])
            cos = (u @ d) / r
            kp = sum((2 * l + 1) * float(sp[c](r)) * eval_legendre(l, cos) for c, l in enumerate(ls)) * w
            tot += float(np.sum(kp * ratio))
        assert abs(E[iw] - tot) < 1e-11 * max(1.0, abs(tot)), (E[iw], tot)


def test_rotation_matrix_known_answers():
    """generateRotationMatrix against the values of Numerics/tests/test_RotationMatrix3D.cpp:25-59."""
    assert np.allclose(orc.rotation_matrix(0.0, 0.0, 0.0), np.eye(3), atol=1e-15)
    r2 = np.array([[-0.459016994374947, 0.842075137094350, -0.283218753550188],
                   [-0.489403985718866, 0.0263932022500210, 0.871657695220709],
                   [0.741476323045772, 0.538714082201795, 0.400000000000000]])
    r3 = np.array([[0.485410196624969, -0.352671151375484, 0.800000000000000],
                   [-0.587785252292473, -0.809016994374947, 9.79717439317882e-17],
                   [0.647213595499958, -0.470228201833979, -0.600000000000000]])
    assert np.allclose(orc.rotation_matrix(0.1, 0.2, 0.3), r2, atol=1e-14)
    assert np.allclose(orc.rotation_matrix(0.9, 0.5, 0.8), r3, atol=1e-14)
    m = orc.rotation_matrix(0.37, 0.61, 0.83)
    assert np.allclose(m @ m.T, np.eye(3), atol=1e-14) and np.isclose(np.linalg.det(m), 1.0)


def test_dmc_local_energy_includes_pseudopotential():
    """DMC's initial local energy = kinetic + Ewald + V_NL with the grid of walker w rotated by the
    first three uniforms of its stream (seed + w)."""
    s = make_system(L=6.0, n_ion=3, n_up=5, n_dn=4, norb=6, M=12, delay_rank=1, seed=7)
    R = make_walkers(s, 4, 8)
    zion = s.N / len(s.ions)
    ls = [0, 1]
    dr, v, v2, _ = channels(ls, seed=2)
    knots, w = icosahedron()
    ref = orc.VMC(s, R, 1, 1)
    _, G, L = ref.evaluate_log()
    base = -0.5 * (np.sum(G * G, axis=(1, 2)) + np.sum(L, axis=1)) + ref.potential(zion).sum(axis=1)
    rot = np.array([orc.rotation_matrix(*orc.rng_uniform(40 + iw, 3)) for iw in range(4)])
    vnl = ref.nlpp_energy(ls, dr, v, v2, 2.0, knots, w, rot)
    assert np.any(np.abs(vnl) > 1e-3)
    o = orc.VMC(s, R, 1, 1)
    o.dmc_set_nlpp(ls, dr, v, v2, 2.0, knots, w)
    o.dmc_init(zion, tau=0.02, capacity=16, seed=40, warmup_steps=2)
    e0 = o.dmc_walker_data()[1]
    assert np.allclose(e0, base + vnl, rtol=1e-12, atol=1e-12)
