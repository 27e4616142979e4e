"""Pin the oracle's optimized long-range breakup (oracle/breakup.py, the reference's default
"opt_breakup" handler) against the reference's known answers and the mathematics.

Known answers: QMCHamiltonians/tests/test_coulomb_pbcAA.cpp:52-163 (CoulombPBCAA with the default
handler): constants -3.1151210154 and energy = Madelung for one charge in a unit cube (as an ion at
the origin and as an electron at (0, 0.5, 0)), constants -1.675229452 and energy -0.9628996199 for
two charges in a BCC arrangement in a cube of side 3.77945227.  The constants lie along the fit's
near-null space (condition number ~1e11) and are checked at the reference's own tolerance (Catch
Approx, 100 float epsilons relative); the energies at 1e-8.

Mathematics: the basis' Hermite conditions at the knots; its Fourier coefficients and r^2
integrals against adaptive quadrature; the fit's misfit; the total energy against the independent
Ewald oracle and against itself at another k-space cutoff.
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import breakup as bk
from oracle import oracle as orc

APPROX = 100 * np.finfo(np.float32).eps
A_BCC = 3.77945227
VMAD_SC = -1.4186487397403098


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.fixture(scope="module")
def unit_cube():
    return bk.Breakup([1.0, 1.0, 1.0])


def test_fit_setup(unit_cube):
    b = unit_cube
    assert b.rc == 0.5 and abs(b.kc - 30.0) < 1e-12 and b.basis.size == 45
    assert b.chi2 < 1e-12 and b.n_lr_shells > 0
    assert len(b.lr_kvectors()) > 0


def test_known_answers_unit_cube(unit_cube):
    v, (sr, lr, consts) = bk.coulomb_aa([1, 1, 1], [[0.0, 0.0, 0.0]], [1.0], brk=unit_cube)
    assert rel(consts, -3.1151210154) < APPROX
    assert rel(v, VMAD_SC) < 1e-7
    v, (sr, lr, consts) = bk.coulomb_aa([1, 1, 1], [[0.0, 0.5, 0.0]], [-1.0], brk=unit_cube)
    assert rel(consts, -3.1151210154) < APPROX and rel(v, -1.418648723) < 1e-8
    assert rel(unit_cube.madelung(), VMAD_SC) < 1e-7


def test_known_answers_bcc():
    v, (sr, lr, consts) = bk.coulomb_aa([A_BCC] * 3, [[0.0, 0.0, 0.0], [1.88972614] * 3], [1.0, 1.0])
    assert rel(consts, -1.675229452) < APPROX and rel(v, -0.9628996199) < 1e-8
    # supercell Madelung constant (CoulombPBCAA.cpp:621-624)
    assert rel(bk.Breakup([A_BCC] * 3).madelung(), VMAD_SC / A_BCC) < 1e-7


def test_basis_hermite_conditions(unit_cube):
    b = unit_cube.basis
    knots = b.delta * np.arange(b.nknots)
    eps = 1e-5 * b.delta
    for n in range(b.size):
        i, kind = divmod(n, 3)
        val = bk.lpqhi_h(b, n, knots)
        # one-sided differences inside [0, r_c]
        lo = np.clip(knots - eps, 0, b.rc)
        hi = np.clip(knots + eps, 0, b.rc)
        f = lambda r: bk.lpqhi_h(b, n, r)
        d1 = (f(hi) - f(lo)) / (hi - lo)
        expect_v = np.where(np.arange(b.nknots) == i, 1.0 if kind == 0 else 0.0, 0.0)
        expect_d = np.where(np.arange(b.nknots) == i, 1.0 / b.delta if kind == 1 else 0.0, 0.0)
        assert np.allclose(val, expect_v, atol=1e-14)
        assert np.allclose(d1 * b.delta, expect_d * b.delta, atol=1e-4)  # one-sided at 0 and r_c
        # support: two knot intervals around knot i
        r = np.linspace(0, b.rc, 301)
        out = (r < b.delta * (i - 1)) | (r > b.delta * (i + 1))
        assert np.all(f(r)[out] == 0)
    # curvature function: second derivative 1 / delta^2 at its own knot (the one-sided cubic terms
    # leave an error of 3 h / delta in the central difference)
    i = 7
    h = 1e-4 * b.delta
    r0 = b.delta * i
    f = lambda r: bk.lpqhi_h(b, 3 * i + 2, np.array([r]))[0]
    assert abs((f(r0 + h) - 2 * f(r0) + f(r0 - h)) / h ** 2 * b.delta ** 2 - 1.0) < 1e-3


def test_fourier_coefficients_and_r2_integrals(unit_cube):
    b = unit_cube.basis
    ks = np.array([2 * np.pi, 30.8, 95.0, 400.0, 3000.0])
    c = bk.lpqhi_c(b, ks)
    for n in (0, 1, 2, 4, 20, 23, 42, 43, 44):
        i = n // 3
        lo, hi = max(0.0, b.delta * (i - 1)), min(b.rc, b.delta * (i + 1))
        for q, k in enumerate(ks):
            f = lambda r: r * np.sin(k * r) * bk.lpqhi_h(b, n, r)
            val = sum(integrate.quad(f, a0, a1, limit=400, epsabs=1e-16, epsrel=1e-13)[0]
                      for a0, a1 in ((lo, b.delta * i), (b.delta * i, hi)) if a1 > a0)
            ref = 4 * np.pi / (k * b.volume) * val
            assert abs(c[n, q] - ref) <= 1e-9 * np.max(np.abs(c[:, q])), (n, k)
        cq = bk.lpqhi_c_quadrature(b, ks[:2])
        for q, k in enumerate(ks[:2]):
            f = lambda r: r * np.sin(k * r) * bk.lpqhi_h(b, n, r)
            val = sum(integrate.quad(f, a0, a1, limit=400, epsabs=1e-16, epsrel=1e-13)[0]
                      for a0, a1 in ((lo, b.delta * i), (b.delta * i, hi)) if a1 > a0)
            assert abs(cq[n, q] - 4 * np.pi / (k * b.volume) * val) <= 1e-13 * np.max(np.abs(cq[:, q]))
        r2 = sum(integrate.quad(lambda r: r * r * bk.lpqhi_h(b, n, r), a0, a1, epsrel=1e-14)[0]
                 for a0, a1 in ((lo, b.delta * i), (b.delta * i, hi)) if a1 > a0)
        assert abs(bk.lpqhi_hintr2(b, n) - r2) < 1e-14


def test_long_range_part_vanishes_above_kc(unit_cube):
    """The fitted V_l's transform is small for k > k_c (what lets the sum stop at k_c) and the
    short-range part goes to zero at r_c."""
    b = unit_cube
    k = np.linspace(b.kc * 1.05, 6000 / b.rc, 2000)
    coulomb = 4 * np.pi / (b.basis.volume * k * k)
    assert np.max(np.abs(b.fk(k)) / coulomb) < 2e-3
    assert abs(b.rvsr(b.rc * (1 - 1e-9))) < 1e-3 and b.rvsr(b.rc) == 0


def test_total_against_ewald_and_cutoff_independence():
    rng = np.random.default_rng(11)
    L = np.array([5.0, 6.0, 7.0])
    r = rng.uniform(0, 1, size=(10, 3)) * L
    q = np.array([1.0, -1.0] * 5)
    v15, p15 = bk.coulomb_aa(L, r, q, dim_cutoff=15.0)
    v25, p25 = bk.coulomb_aa(L, r, q, dim_cutoff=25.0)
    ve, _ = orc.coulomb_aa(L, r, q, dim_cutoff=45.0)  # Ewald, converged to 1e-10
    assert abs(v15 - ve) < 1e-6 * abs(ve) and abs(v25 - ve) < 1e-7 * abs(ve)
    assert abs(p15[1] - p25[1]) > 1e-3  # the split moves


def test_spline_energy_matches_exact_short_range():
    """The device tabulates r V_s on the reference's 1001-point cubic spline; the oracle's spline
    evaluation agrees with the exact short-range function to the spline's accuracy."""
    rng = np.random.default_rng(12)
    L = np.array([6.0, 6.5, 7.0])
    r = rng.uniform(0, 1, size=(12, 3)) * L
    q = rng.choice([-1.0, 2.0], size=12)
    brk = bk.Breakup(L)
    ve, _ = bk.coulomb_aa(L, r, q, brk=brk)
    vs, _ = bk.coulomb_aa(L, r, q, brk=brk, spline=True)
    assert abs(ve - vs) < 1e-6 * max(1.0, abs(ve))


def test_kcontainer_and_structure_factor_known_answers():
    """The lattice k list and the structure factor: 78 shells at k_c = 12 in a cube of side 5
    (Particle/LongRange/tests/test_lrhandler.cpp:30-52), 263786 vectors at k_c = 50 and the sum
    over them of each species' rho_k for 3 + 1 particles (test_StructFact.cpp:24-70, checked
    there at 5e-5)."""
    _, shell_k2, _ = bk.lattice_kpoints([5.0, 5.0, 5.0], 12.0)
    assert len(shell_k2) == 78
    K, _, _ = bk.lattice_kpoints([5.0, 5.0, 5.0], 50.0)
    assert len(K) == 263786
    R = np.array([[0.0, 1.0, 2.0], [1.0, 0.2, 3.0], [0.3, 4.0, 1.4], [3.2, 4.7, 0.7]])
    sums = [np.sum(np.exp(1j * (K @ R[:3].T))), np.sum(np.exp(1j * (K @ R[3:].T)))]
    for got, want in zip(sums, [-125.80618630936, 68.199075127271]):
        assert abs(got - want) < 1e-9 * abs(want)
