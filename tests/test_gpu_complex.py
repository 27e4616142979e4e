"""Complex (C2C) orbitals on the HIP path against the CPU FP64 oracle (same seeded inputs).

The oracle's complex orbitals and complex determinant engine are pinned by
tests/test_oracle_complex.py (composition with the real evaluator, finite differences, Bloch
periodicity, complex LU against numpy, ratios against recomputation, the real limit).  Here the
device's C2C orbitals, complex determinant recompute (rocSOLVER zgetrf/zgetri), delayed updates,
flush and mw_evaluateGL are compared with it: values, gradients, Laplacians and local energies
within 1e-10 relative (ratios relative to max(|ratio|, 1e-2)), the phase of psi modulo 2 pi,
and identical accept/reject decisions on the oracle's replay of the device trajectory.  The trace
holds |ratio| for complex orbitals.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import System, jastrow_params, make_complex_system, make_walkers

from test_gpu_parity import REL, RATIO_FLOOR, check_vmc, quadrature_positions, relerr

pytestmark = pytest.mark.gpu


def complex_small(n_up=5, n_dn=4, k=1, seed=3, norb=6, nion=7):
    rng = np.random.default_rng(seed)
    L = np.array([7.0, 8.0, 9.5])
    M = (10, 12, 11)
    p = jastrow_params(8)
    return System(L=L, ions=rng.uniform(0, 1, size=(nion, 3)) * L, n_up=n_up, n_dn=n_dn, M=M, norb=norb,
                  coefs=rng.uniform(-1, 1, size=M + (norb, 2)) + 0.3,
                  j1_params=p * 0.7, j1_rcut=3.4, j1_cusp=0.0, j2_uu=p, j2_ud=p * 1.3, j2_rcut=3.4,
                  j2_cusp_uu=-0.25, j2_cusp_ud=-0.5, delay_rank=k, kpoints=rng.uniform(-1, 1, size=(norb, 3)))


def ctx(system, R, wpc):
    c = qmi.Context(0)
    c.system_create(system)
    c.walkers_load(R, wpc)
    assert c.cplx
    return c


def phase_err(a, b):
    return float(np.max(np.abs(np.angle(np.exp(1j * (np.imag(a) - np.imag(b)))))))


def test_c2c_spo_vgl_and_det_ratios():
    s = complex_small(norb=37)
    rng = np.random.default_rng(11)
    R = rng.uniform(0, 1, size=(4, s.N, 3)) * s.L
    c = ctx(s, R, 1)
    pos = rng.uniform(-10, 20, size=(4, 3))
    out = c.mw_evaluateVGL(pos, 37)
    for w in range(4):
        ref = orc.spo_vgl_c(s.L, s.coefs, s.kpoints, pos[w])
        assert relerr(out[w], ref, floor=1e-3) < 1e-12
    nvp = 5
    vpos = rng.uniform(-10, 20, size=(4, nvp, 3))
    inv = rng.standard_normal((4, 37)) + 1j * rng.standard_normal((4, 37))
    rat = c.mw_evaluateDetRatios(vpos, inv)
    ref = np.array([[orc.spo_vgl_c(s.L, s.coefs, s.kpoints, vpos[w, v])[0] @ inv[w] for v in range(nvp)]
                    for w in range(4)])
    assert relerr(rat, ref, floor=1.0) < 1e-12


@pytest.mark.parametrize("k", [1, 3])
def test_c2c_evaluate_log_and_protocol(k):
    """mw_evaluateLog, then evalGrad / makeMove / calcRatioGrad / accept_rejectMove (host
    decisions) / completeUpdates / evaluateGL, all complex, against the oracle."""
    s = complex_small(k=k)
    nw = 3
    rng = np.random.default_rng(21)
    R = rng.uniform(0, 1, size=(nw, s.N, 3)) * s.L
    c = ctx(s, R, 1)
    lp, G, L = c.mw_evaluateLog()
    o = orc.VMC(s, R, 1, 1)
    lp_o, G_o, L_o = o.evaluate_log()
    assert relerr(lp.real, lp_o.real, floor=1.0) < REL and phase_err(lp, lp_o) < REL
    assert relerr(G, G_o, floor=1.0) < REL and relerr(L, L_o, floor=1.0) < REL
    acc_pattern = [1, 0, 1, 1, 0, 1, 1, 1, 0]
    for sweep in range(2):
        for iat in range(s.N):
            g = c.mw_evalGrad(iat)
            for w in range(nw):
                assert relerr(g[w], o.eval_grad(w, iat), floor=1.0) < REL
            displ = rng.normal(size=(nw, 3)) * 0.5
            c.mw_makeMove(iat, displ)
            r, gn = c.mw_calcRatioGrad(iat)
            acc = np.array([acc_pattern[(iat + w + sweep) % len(acc_pattern)] for w in range(nw)])
            for w in range(nw):
                Rw = o.get()["R"][w]
                ro, go = o.ratio_grad(w, iat, Rw[iat] + displ[w])
                assert relerr(r[w], ro, floor=RATIO_FLOOR) < REL
                assert relerr(gn[w], go, floor=1.0) < REL
                o.accept_reject(w, iat, bool(acc[w]))
            c.mw_accept_rejectMove(iat, acc)
        c.mw_completeUpdates()
        lp, G, L, ek = c.mw_evaluateGL(fromscratch=False)
        for w in range(nw):
            lpo, Go, Lo = o.evaluate_gl(w)
            assert abs(lp[w].real - lpo.real) < REL * max(1.0, abs(lpo.real)) and phase_err(lp[w], lpo) < REL
            assert relerr(G[w], Go, floor=1.0) < REL and relerr(L[w], Lo, floor=1.0) < REL
            ke = -0.5 * np.real(np.sum(Go * Go) + np.sum(Lo))
            assert abs(ek[w] - ke) < REL * max(1.0, abs(ke))
    assert np.allclose(c.walkers_get(), o.get()["R"], rtol=0, atol=1e-12)


def vmc_parity_c(system, R, wpc, steps, sub_steps, tau, recompute_period, seed):
    c = ctx(system, R, wpc)
    c.mw_evaluateLog()
    E, tr = c.vmc_run(steps, sub_steps, tau, recompute_period, seed, trace=True)
    o = orc.VMC(system, R, wpc, seed)
    o.evaluate_log()
    Eo, tro = o.replay(tr, sub_steps, tau, recompute_period)
    return c, o, E, tr, Eo, tro


@pytest.mark.parametrize("k,wpc", [(1, 1), (4, 1), (8, 3)])
def test_c2c_vmc_delayed(k, wpc):
    """VMC with complex orbitals: the fused move kernel (one walker per crowd) and the two-kernel
    crowd path, rank-k delayed updates with the complex flush, sub-steps and a recompute."""
    s = complex_small(n_up=12, n_dn=10, k=k, norb=14, seed=5)
    R = np.random.default_rng(6).uniform(0, 1, size=(6, s.N, 3)) * s.L
    c, o, E, tr, Eo, tro = vmc_parity_c(s, R, wpc, steps=3, sub_steps=2, tau=0.4, recompute_period=2, seed=77)
    check_vmc(c, o, E, tr, Eo, tro)
    lp = c.mw_evaluateGL(fromscratch=False)[0]
    lpo = o.get()["logpsi"]
    assert relerr(lp.real, lpo.real, floor=1.0) < REL and phase_err(lp, lpo) < 1e-9


@pytest.mark.parametrize("n_up,n_dn,k", [(128, 128, 32), (300, 260, 16)])
def test_c2c_vmc_large(n_up, n_dn, k):
    """Twisted complex orbitals at production determinant sizes (a quarter-sized cell of the
    1024-electron density), including determinants beyond 256 orbitals (two orbitals per thread).
    Ratio moduli (relative to max(|ratio|, 1e-2)) and gradients are within 1e-10 of the oracle and
    of its long double determinant shadow (the exact reference), decisions identical, proposals
    and energies within 1e-10."""
    s = make_complex_system(L=14.0, n_ion=16, n_up=n_up, n_dn=n_dn, norb=max(n_up, n_dn), M=12,
                            delay_rank=k, seed=41)
    R = make_walkers(s, 2, 42)
    c = ctx(s, R, 1)
    c.mw_evaluateLog()
    E, tr = c.vmc_run(1, 1, 0.3, 0, 80, trace=True)
    o = orc.VMC(s, R, 1, 80)
    o.evaluate_log()
    Eo, tro, ex = o.replay(tr, 1, 0.3, 0, exact=True)
    dev, own = relerr(tr[..., 0], ex[..., 0], floor=RATIO_FLOOR), relerr(tro[..., 0], ex[..., 0], floor=RATIO_FLOOR)
    gdev, gown = relerr(tr[..., 5:8], ex[..., 5:8], floor=1.0), relerr(tro[..., 5:8], ex[..., 5:8], floor=1.0)
    print(f"vs exact: |ratio| device {dev:.3e}, oracle {own:.3e}; gradient device {gdev:.3e}, oracle {gown:.3e}")
    assert dev < REL and gdev < REL
    check_vmc(c, o, E, tr, Eo, tro, ratio_floor=RATIO_FLOOR)


def test_c2c_evaluate_ratios_virtual_moves():
    """mw_evaluateRatios with complex orbitals (complex ratios) after a delayed-update VMC run."""
    s = complex_small(n_up=12, n_dn=10, k=4, norb=14, seed=5)
    nw = 4
    R = np.random.default_rng(6).uniform(0, 1, size=(nw, s.N, 3)) * s.L
    c, o, E, tr, Eo, tro = vmc_parity_c(s, R, 1, steps=2, sub_steps=1, tau=0.4, recompute_period=0, seed=82)
    Rc = c.walkers_get()
    rng = np.random.default_rng(8)
    worst = 0.0
    for iat in (1, 11, 20):
        vp = quadrature_positions(Rc, iat, s.ions, rng)
        rat = c.mw_evaluateRatios(iat, vp)
        ref = np.array([[o.ratio_grad(w, iat, vp[w, v])[0] for v in range(vp.shape[1])] for w in range(nw)])
        worst = max(worst, relerr(rat, ref, floor=RATIO_FLOOR))
    print("complex virtual-move ratio error", worst)
    assert worst < REL
