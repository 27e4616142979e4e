"""Pin the CPU oracle to the reference's own known-answer tests.

Every expected number comes from tests/golden/reference_kats.json, which cites
the reference test file:line it was transcribed from.  The reference's Catch2
Approx() tolerance is about 1.2e-5 relative; the values carry ~10 digits, so
the checks below use 2e-9 relative where the value has 10 significant digits.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd.system import System

REL = 2e-9


def close(a, b, rel=REL, abs_=1e-12):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.all(np.abs(a - b) <= rel * np.abs(b) + abs_)


# ---------------------------------------------------------------------------------------- RNG
def test_stdrandom_seed13(kats):
    k = kats["rng_stdrandom_seed13"]
    u = orc.rng_uniform(k["seed"], 3)
    assert close(u, k["expected"])


def test_mt19937_standard_answer(kats):
    k = kats["mt19937_10000th"]
    u = orc.rng_uniform(k["seed"], 10000)
    assert int(round(u[-1] * 2.0 ** 32)) == k["expected"]


@pytest.mark.parametrize("case", range(5))
def test_box_muller_fake_random(kats, case):
    """assignGaussRand over the reference's constant generator: pairs (r cos t, r sin t), the odd
    tail's cosine, the guard against u = 1, and nothing written past n."""
    c = kats["gauss_fake_random"]["cases"][case]
    g = orc.rng_gauss_const(c["value"], c["n"])  # n outputs and one slot that must stay zero
    e = c["expected"]  # the reference's array: n outputs, then (for n < len) the untouched slot
    # the reference checks these with Catch's default Approx (1.19e-5 relative); its listed values
    # carry ~1e-10 relative noise around the closed form sqrt(2 ln 2) (cos pi, sin pi) at u = 0.5;
    # sin(2 pi) of the u = 1 case is -2.4e-16 in FP64, hence the absolute floor
    assert close(g[:len(e)], e, rel=1.2e-5, abs_=1e-14)
    assert g[c["n"]] == 0.0
    if c["value"] == 0.5:
        assert abs(g[0] + np.sqrt(2 * np.log(2))) < 1e-15


def test_box_muller_distribution():
    """The Gaussians of the mt19937 stream are standard normal (moments and the tail fraction of
    2e5 draws), so no term of the transform is dropped for general uniforms."""
    g = orc.rng_gauss(2024, 200001)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1) < 0.01
    assert abs(np.mean(np.abs(g) > 1.959964) - 0.05) < 0.003
    assert abs(np.mean(g ** 3)) < 0.02 and abs(np.mean(g ** 4) - 3) < 0.05


# ---------------------------------------------------------------------------------------- driver
def test_drift_unr(kats):
    k = kats["drift_unr"]
    tom = k["tau"] / k["mass"]
    for i in range(k["nx"]):
        gx = -k["xtot"] / 2 + i * k["xtot"] / k["nx"]
        d = orc.get_drift(tom, [gx, 0.0, 0.0])
        f = gx * gx * tom
        ref = (-1 + np.sqrt(1 + 2 * f)) / f * gx * tom if f > 0 else tom * gx
        assert abs(d[0] - ref) <= 1e-12 * max(1, abs(ref)) and d[1] == 0 and d[2] == 0


def test_tau_params(kats):
    k = kats["tau_params"]
    t = orc.tau_params(k["tau"], k["invmass"])
    assert close(t, [k["tauovermass"], k["oneover2tau"], k["sqrttau"]], rel=1e-15)


# The reference checks driver positions with Catch's default Approx (1.19e-5 relative); the digits
# it lists differ between its own tests by ~3e-11 (test_vmc.cpp vs test_dmc.cpp for the same move),
# so the oracle is held to that tolerance and, beyond it, to the closed form of the move:
# R1 = R0 + drift + sqrt(tau) xi, xi the Box-Muller pairs of u = 0.5 filled electron by electron.
APPROX = 1.2e-5


def _closed_form_move(R0, tau, grad):
    r = np.sqrt(2 * np.log(2))
    xi = np.array([r * np.cos(np.pi), r * np.sin(np.pi)] * 3).reshape(2, 3)
    g = np.asarray(grad, dtype=float)
    v2 = g @ g
    drift = g * ((-1 + np.sqrt(1 + 2 * tau * v2)) / v2 if v2 > 0 else tau)
    return np.asarray(R0) + drift + np.sqrt(tau) * xi


def _two_electron_open_cell(R0):
    """The reference driver tests' setup: one species of two electrons, one ion at the origin,
    an open cell, and no wave-function components besides the test-only one."""
    return System(L=np.zeros(3), ions=np.zeros((1, 3)), n_up=2, n_dn=0, M=(0, 0, 0), norb=0, coefs=None,
                  j1_params=None, j1_rcut=1.0, j1_cusp=0.0, j2_uu=None, j2_ud=None, j2_rcut=1.0,
                  j2_cusp_uu=0.0, j2_cusp_ud=0.0, delay_rank=1), np.array([R0], dtype=float)


def test_vmc_pbyp_constant_orbital(kats):
    """VMC drift-diffusion step of test_vmc.cpp: ConstantOrbital (no drift), FakeRandom 0.5."""
    k = kats["vmc_pbyp_constant_orbital"]
    s, R = _two_electron_open_cell(k["R0"])
    o = orc.VMC(s, R, 1, 1)
    o.test_orbital("constant")
    o.fake_rng(0.5)
    o.evaluate_log()
    o.sweep(k["tau"])
    st = o.get()
    assert np.allclose(st["R"][0], k["R1"], rtol=APPROX, atol=1e-14)
    assert np.allclose(st["R"][0], _closed_form_move(k["R0"], k["tau"], [0, 0, 0]), rtol=1e-15, atol=1e-15)
    assert (st["accepted"], st["rejected"]) == (k["accepted"], k["rejected"])


def test_dmc_pbyp_constant_orbital(kats):
    """DMC step of test_dmc.cpp: accepted moves, then a negative ratio (a node crossing of a real
    wave function) rejects both moves before the draw."""
    k = kats["dmc_pbyp_constant_orbital"]
    s, R = _two_electron_open_cell(k["R0"])
    o = orc.VMC(s, R, 1, 1)
    o.test_orbital("constant")
    o.fake_rng(0.5)
    o.evaluate_log()
    o.sweep(k["tau"], dmc=True)
    st = o.get()
    assert np.allclose(st["R"][0], k["R1"], rtol=APPROX, atol=1e-14)
    assert np.allclose(st["R"][0], _closed_form_move(k["R0"], k["tau"], [0, 0, 0]), rtol=1e-15, atol=1e-15)
    assert (st["accepted"], st["rejected"]) == (k["after_first"]["accepted"], k["after_first"]["rejected"])
    o.test_orbital("constant", fake_ratio=-1.0)
    o.sweep(k["tau"], dmc=True)
    st = o.get()
    assert np.allclose(st["R"][0], k["R1"], rtol=APPROX, atol=1e-14)
    assert (st["accepted"], st["rejected"]) == (k["after_second_with_ratio_m1"]["accepted"],
                                                k["after_second_with_ratio_m1"]["rejected"])


def test_dmc_pbyp_linear_orbital(kats):
    """DMC step of test_dmc.cpp with LinearOrbital: the UNR drift of the gradient (1, 2, 3), the
    Green's function ratio with no gradient at the proposed position, FakeRandom 0.5."""
    k = kats["dmc_pbyp_linear_orbital"]
    s, R = _two_electron_open_cell(k["R0"])
    o = orc.VMC(s, R, 1, 1)
    o.test_orbital("linear")
    o.fake_rng(0.5)
    o.evaluate_log()
    o.sweep(k["tau"], dmc=True)
    st = o.get()
    assert np.allclose(st["R"][0], k["R1"], rtol=APPROX, atol=1e-14)
    assert np.allclose(st["R"][0], _closed_form_move(k["R0"], k["tau"], [1, 2, 3]), rtol=1e-14, atol=1e-15)
    assert (st["accepted"], st["rejected"]) == (k["accepted"], k["rejected"])


# ---------------------------------------------------------------------------------------- spline
def test_symtrace(kats):
    k = kats["symtrace"]
    assert abs(orc.symtrace(k["h"], k["gg"]) - k["expected"]) < 1e-12


def test_prefactors(kats):
    k = kats["prefactors"]
    a, _, _ = orc.prefactors(0.1)
    assert close(a, k["tx_0.1_a"], rel=5e-6)
    a, da, d2a = orc.prefactors(0.8)
    assert close(a, k["tx_0.8_a"], rel=5e-6)
    assert close(da, k["tx_0.8_da"], rel=1e-12)
    assert close(d2a, k["tx_0.8_d2a"], rel=1e-12)


def periodic_interp_coefs(data):
    """Solve (P[i-1] + 4 P[i] + P[i+1]) / 6 = data[i] along each periodic axis (einspline's
    interpolating periodic cubic B-spline, bspline_create.c:960-1034) via FFT."""
    P = data.astype(np.complex128)
    for ax in range(data.ndim):
        M = data.shape[ax]
        m = np.arange(M)
        denom = (4.0 + 2.0 * np.cos(2 * np.pi * m / M)) / 6.0
        shape = [1] * data.ndim
        shape[ax] = M
        P = np.fft.ifft(np.fft.fft(P, axis=ax) / denom.reshape(shape), axis=ax)
    return P.real


def test_multi_spline_5(kats):
    k = kats["multi_spline_5"]
    N = k["grid"]
    x = np.arange(N) / N
    data = (np.sin(2 * np.pi * x)[:, None, None] + np.sin(3 * 2 * np.pi * x)[None, :, None]
            + np.sin(4 * 2 * np.pi * x)[None, None, :])
    P = periodic_interp_coefs(data)[..., None]
    out = orc.spline_vgh_raw(P, k["pos0"])
    assert abs(out[0, 0] - k["v0"]) < 1e-10
    assert close(out[1:4, 0], k["g0"])
    assert np.all(np.abs(out[4:10, 0]) < 1e-8)
    out = orc.spline_vgh_raw(P, k["pos1"])
    assert close(out[0, 0], k["v1"])
    assert close(out[1:4, 0], k["g1"])
    h = out[4:10, 0]
    for i in (0, 3, 5):
        assert close(h[i], k["h1"][i])
    for i in (1, 2, 4):  # analytically zero, ~1e-9 from the interpolation solve
        assert abs(h[i]) < 1e-8
    assert close(h[0] + h[3] + h[5], k["lap1"])


# ---------------------------------------------------------------------------------------- Jastrow
def open_system(n_up, n_dn, ions=None, j1=None, j2=None):
    ions = np.zeros((0, 3)) if ions is None else np.asarray(ions, dtype=float)
    j1 = j1 or dict(params=None, rcut=0.0, cusp=0.0)
    j2 = j2 or dict(uu=None, ud=None, rcut=0.0, cusp_uu=0.0, cusp_ud=0.0)
    return System(L=np.zeros(3), ions=ions, n_up=n_up, n_dn=n_dn, M=(0, 0, 0), norb=0, coefs=None,
                  j1_params=j1["params"], j1_rcut=j1["rcut"], j1_cusp=j1["cusp"],
                  j2_uu=j2["uu"], j2_ud=j2["ud"], j2_rcut=j2["rcut"], j2_cusp_uu=j2["cusp_uu"],
                  j2_cusp_ud=j2["cusp_ud"], delay_rank=1)


def test_j2_functor_table(kats):
    k = kats["j2_bspline"]
    for r, u, du, ddu in k["table_r_u_du_ddu"]:
        v = orc.functor_vgl(k["params_ud"], k["rcut"], k["cusp_ud"], r)
        assert close(v, [u, du, ddu], rel=REL, abs_=2e-11), (r, v, (u, du, ddu))


def test_j2_log_ratio_accept(kats):
    k = kats["j2_bspline"]
    p = np.array(k["params_ud"])
    sysj = open_system(1, 1, j2=dict(uu=p, ud=p, rcut=k["rcut"], cusp_uu=-0.25, cusp_ud=k["cusp_ud"]))
    R = np.array(k["elec"])[None]
    vmc = orc.VMC(sysj, R, 1, 1)
    logpsi, G, L = vmc.evaluate_log()
    assert close(logpsi[0], k["logpsi"])
    KE = -0.5 * (np.sum(G[0] * G[0]) + np.sum(L[0]))
    assert close(KE, k["KE"])
    r0, _ = vmc.ratio_grad(0, 0, k["newpos"])
    assert close(r0, k["ratio_e0"])
    vmc.accept_reject(0, 0, False)
    r1, _ = vmc.ratio_grad(0, 1, k["newpos"])
    assert close(r1, k["ratio_e1"])
    vmc.accept_reject(0, 1, True)
    lp, _, _ = vmc.evaluate_gl(0)
    assert close(lp, k["logpsi_after_accept_e1"])


def test_j1_log_ratio_accept(kats):
    k = kats["j1_bspline"]
    sysj = open_system(1, 1, ions=[k["ion"]], j1=dict(params=np.array(k["params"]), rcut=k["rcut"], cusp=k["cusp"]))
    vmc = orc.VMC(sysj, np.array(k["elec"])[None], 1, 1)
    logpsi, _, _ = vmc.evaluate_log()
    assert close(logpsi[0], k["logpsi"])
    r0, _ = vmc.ratio_grad(0, 0, k["newpos"])
    assert close(r0, k["ratio_e0"])
    vmc.accept_reject(0, 0, False)
    r1, _ = vmc.ratio_grad(0, 1, k["newpos"])
    assert close(r1, k["ratio_e1"])
    vmc.accept_reject(0, 1, True)
    lp, _, _ = vmc.evaluate_gl(0)
    assert close(lp, k["logpsi_after_accept_e1"])


# ---------------------------------------------------------------------------------------- determinant
def test_ddb_first(kats):
    k = kats["ddb_first"]
    a = np.array(k["a"])
    n = 3
    det = orc.Det(1, n, 1)
    psiM = a.T.copy()  # FakeSPO::evaluate_notranspose: logdet(j, i) = a(i, j)
    gv = np.array(kats["ddb_4x4"]["gv"])
    dpsiM = np.array([[gv[j] + i for j in range(n)] for i in range(n)])
    assert det.recompute(0, psiM, dpsiM, np.zeros((n, n))) == 0
    ainv, ld0 = det.get(0)
    assert close(ainv, k["inv0"])

    def phi_of(v):
        out = np.zeros((5, n))
        out[0] = v
        return out

    r, _ = det.ratio_grad(0, phi_of(k["v"]))
    assert close(r[0], k["ratio_v_row0"], rel=1e-11)
    det.accept_reject(0, [1])
    ainv, _ = det.get(0)
    assert close(ainv, k["inv1"])
    r, _ = det.ratio_grad(0, phi_of(a[0]))
    assert close(r[0], k["ratio_a0_row0"])
    det.accept_reject(0, [0])
    r, _ = det.ratio_grad(1, phi_of(a[1]))
    assert close(r[0], k["ratio_a1_row1"])
    det.accept_reject(1, [1])
    _, ld = det.get(0)
    assert close(ld.real, k["log_after"])


@pytest.mark.parametrize("k_delay", [1, 2, 3])
def test_ddb_4x4_sequences(kats, k_delay):
    """test_DiracDeterminantBatched second/delayed_update: three successive accepted row
    replacements, each ratio equal to det(new)/det(old) and the final inverse equal to a fresh
    inversion; run through the batched API with two walkers in lock-step."""
    k = kats["ddb_4x4"]
    a2, v2 = np.array(k["a2"]), np.array(k["v2"])
    n = 4
    det = orc.Det(2, n, k_delay)
    for iw in range(2):
        det.recompute(iw, a2.T.copy(), np.zeros((n, n, 3)), np.zeros((n, n)))
    cur = a2.T.copy()  # psiM rows = electrons
    for row in range(3):
        phi = np.zeros((2, 5, n))
        phi[:, 0] = v2[row]
        if row > 0:
            det.eval_grad(row)
        r, _ = det.ratio_grad(row, phi)
        new = cur.copy()
        new[row] = v2[row]
        expect = np.linalg.det(new) / np.linalg.det(cur)
        assert close(r, [expect, expect], rel=1e-12)
        det.accept_reject(row, [1, 1])
        cur = new
    det.complete()
    for iw in range(2):
        ainv, ld = det.get(iw)
        assert close(ainv, np.linalg.inv(cur.T), rel=1e-12, abs_=1e-12)
        s, lda = np.linalg.slogdet(cur)
        assert abs(ld.real - lda) < 1e-12
        assert (np.round(ld.imag / np.pi) % 2 == 0) == (s > 0)


# ---------------------------------------------------------------------------------------- distance
# ---------------------------------------------------------------------------------------- LU inverse
def _wrap_phase(x):
    return (x + np.pi) % (2 * np.pi) - np.pi


@pytest.mark.parametrize("name", ["dirac_matrix_real", "dirac_matrix_testmatrix1"])
def test_dirac_matrix_real(kats, name):
    """DiracMatrix::invert_transpose known answers: inverse and log det of 3x3 and 4x4 matrices that
    need row interchanges (the 4x4 determinant is negative twice over: phase 2 pi)."""
    k = kats[name]
    a = np.array(k["a"], dtype=float)
    inv, logdet, info = orc.invert_transpose(np.ascontiguousarray(a.T))
    assert info == 0
    assert close(inv, k["inv"], rel=1e-9, abs_=1e-10)
    lv = np.atleast_1d(k["log_value"])
    assert abs(logdet.real - lv[0]) < 1e-5 * abs(lv[0])
    if lv.size > 1:
        assert abs(_wrap_phase(logdet.imag - lv[1])) < 1e-5


def test_dirac_matrix_complex(kats):
    k = kats["dirac_matrix_complex"]
    a = np.array(k["a_re"]) + 1j * np.array(k["a_im"])
    inv, logdet, info = orc.invert_transpose_c(a)
    assert info == 0
    ref = np.array(k["inv_re"]) + 1j * np.array(k["inv_im"])
    assert np.max(np.abs(inv - ref)) < 1e-14
    assert abs(logdet.real - k["log_value"][0]) < 1e-13
    assert abs(_wrap_phase(logdet.imag - k["log_value"][1])) < 1e-13


def test_distance_pbc(kats):
    k = kats["distance_pbc"]
    e = np.array(k["elec"])
    e0 = e[0] + np.array(k["shift_e0"]) + np.array(k["disp"])
    r, dr = orc.min_image([k["L"]] * 3, e0[None], e[1])
    assert close(r[0], k["dist_row1_col0"])
    assert close(dr[0], k["displ_row1_col0"], rel=1e-12, abs_=1e-12)
    # moving a particle by a full lattice vector changes no distance
    r2, dr2 = orc.min_image([k["L"]] * 3, (e0 + np.array([6.0, -6.0, 12.0]))[None], e[1])
    assert abs(r2[0] - r[0]) < 1e-12
