"""Learned per-electron proposal on the device against the CPU reference (oracle/learned.py) and
the FP64 oracle.

* network forward (BF16 operands, FP32 accumulation) and parameter gradients vs the reference;
* features vs the reference definition;
* a zero output layer is plain drift-diffusion: trajectory replay against the plain oracle with
  identical decisions and FP64 ratios;
* with a random network, the acceptance probability and the training objective recomputed from
  the oracle's ratios and the traced proposal;
* exact sampling: one electron, random network, histogram of x against the exact marginal of
  |psi|^2 from the oracle;
* training raises the objective (acceptance-weighted squared jump) over the untrained proposal.
"""
import dataclasses

import numpy as np
import pytest

from oracle import learned as lp
from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import System, jastrow_params, make_system, make_walkers

pytestmark = pytest.mark.gpu


def flat(P):
    return np.concatenate([P[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")]).astype(np.float32)


def unflat(v):
    P, off = {}, 0
    for k, shp in (("W1", (128, 68)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, 128)),
                   ("b3", (128,)), ("W4", (4, 128)), ("b4", (4,))):
        n = int(np.prod(shp))
        P[k] = np.asarray(v[off:off + n], dtype=np.float64).reshape(shp)
        off += n
    return P


def ctx_for(s, R, precision="fp64"):
    c = qmi.Context(0, precision=precision)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.mw_evaluateLog()
    return c


def small_system(seed=5, n_up=12, n_dn=10, norb=14, k=4):
    rng = np.random.default_rng(seed)
    L = np.array([7.0, 8.0, 9.5])
    M = (10, 12, 11)
    p = jastrow_params(8)
    return System(L=L, ions=rng.uniform(0, 1, size=(7, 3)) * L, n_up=n_up, n_dn=n_dn, M=M, norb=norb,
                  coefs=rng.uniform(-1, 1, size=M + (norb,)) + 0.3, j1_params=p * 0.7, j1_rcut=3.4, j1_cusp=0.0,
                  j2_uu=p, j2_ud=p * 1.3, j2_rcut=3.4, j2_cusp_uu=-0.25, j2_cusp_ud=-0.5, delay_rank=k)


def test_mlp_forward_and_gradient():
    rng = np.random.default_rng(7)
    s = small_system()
    c = ctx_for(s, rng.uniform(0, 1, (2, s.N, 3)) * s.L)
    P = lp.init_params(rng, zero_head=False)
    c.mlp_params(flat(P))
    X = rng.standard_normal((37, 68)) * 2.0
    out = c.mlp_forward(X)
    ref, cache = lp.forward(unflat(flat(P)), X.astype(np.float32).astype(np.float64), emulate=True)
    err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    print("forward max error / max |out|:", err)
    assert err < 2e-2
    dout = rng.standard_normal((37, 4))
    g = c.mlp_gradient(X, dout)
    G = lp.backward(unflat(flat(P)), cache, dout)
    gref = flat(G)
    for name, sl in (("W1", slice(0, 8704)), ("W2", slice(8832, 25216)), ("W4", slice(41856, 42368)),
                     ("all", slice(0, 42372))):
        e = np.linalg.norm(g[sl] - gref[sl]) / np.linalg.norm(gref[sl])
        print(f"gradient {name}: relative error {e:.2e}")
        assert e < 3e-2, name


def test_features_match_reference():
    rng = np.random.default_rng(8)
    s = small_system(n_up=20, n_dn=18, norb=22)
    R = rng.uniform(0, 1, (3, s.N, 3)) * s.L
    c = ctx_for(s, R)
    iat = 5
    pos = R[:, iat] + rng.normal(0, 0.4, (3, 3))
    drift = rng.normal(0, 0.2, (3, 3))
    x = c.proposal_features(iat, pos, drift, 0.3)
    for w in range(3):
        ref = lp.features(s.L, s.ions, R[w], iat, pos[w], drift[w], 0.3)
        assert np.max(np.abs(x[w] - ref)) < 1e-5 * max(1.0, np.max(np.abs(ref)))


def replay_positions(R0, tr):
    """positions r of the moving electron before each traced move [steps][sub][N][nw][3]"""
    R = R0.copy()
    steps, sub, N, nw = tr.shape[:4]
    r = np.empty((steps, sub, N, nw, 3))
    for a in range(steps):
        for b in range(sub):
            for i in range(N):
                r[a, b, i] = R[:, i]
                acc = tr[a, b, i, :, 1] != 0
                R[acc, i] = tr[a, b, i, acc, 2:5]
    return r


# FP64 engine with the full table and with the table-only narrowings (FP32, BF16 storage): the oracle
# runs on the table as the device stores it, so FP64 parity holds in all three
ENGINE_FP64 = ["fp64", "t32", "tbf16"]


@pytest.mark.parametrize("precision", ENGINE_FP64)
def test_zero_head_is_plain_drift_diffusion(precision):
    s = small_system()
    R = np.random.default_rng(6).uniform(0, 1, size=(6, s.N, 3)) * s.L
    c = ctx_for(s, R, precision)
    c.proposal_config(learned=True, train=False, init_seed=99)
    E, tr, lt = c.vmc_run_learned(3, 2, 0.4, 2, 77, trace=True)
    assert np.all(lt[..., 0:4] == 0) and np.all(lt[..., 4:8] == 0)
    s = dataclasses.replace(s, coefs=orc.round_table(s.coefs, precision))
    o = orc.VMC(s, R, 1, 77)
    o.evaluate_log()
    Eo, tro = o.replay(tr, 2, 0.4, 2)
    nmis = int(np.sum(tr[..., 1] != tro[..., 1]))
    rerr = np.max(np.abs(tr[..., 0] - tro[..., 0]) / np.maximum(np.abs(tro[..., 0]), 1e-2))
    eerr = np.max(np.abs(E - Eo) / np.maximum(np.abs(Eo), 1.0))
    print(f"decisions differing {nmis}, ratio {rerr:.2e}, energy {eerr:.2e}")
    assert nmis == 0 and rerr < 1e-10 and eerr < 1e-10
    assert np.array_equal(c.walkers_get(), o.get()["R"])


@pytest.mark.parametrize("precision", ENGINE_FP64)
def test_random_network_acceptance_and_objective(precision):
    rng = np.random.default_rng(9)
    s = small_system()
    R = rng.uniform(0, 1, size=(6, s.N, 3)) * s.L
    c = ctx_for(s, R, precision)
    P = lp.init_params(rng, zero_head=False, scale=0.5)
    c.mlp_params(flat(P))
    c.proposal_config(learned=True, train=False)
    tau = 0.4
    E, tr, lt = c.vmc_run_learned(2, 1, tau, 0, 78, trace=True)
    s = dataclasses.replace(s, coefs=orc.round_table(s.coefs, precision))
    o = orc.VMC(s, R, 1, 78)
    o.evaluate_log()
    Eo, tro = o.replay(tr, 1, tau, 0)
    assert relerr_ratio(tr, tro) < 1e-10
    r = replay_positions(R, tr)
    rp = tr[..., 2:5]
    v = np.apply_along_axis(lambda g: lp.unr_drift(tau, g), -1, tro[..., 5:8])
    mu, s_f, mu_r, s_r, vr = lt[..., 0:3], lt[..., 3], lt[..., 4:7], lt[..., 7], lt[..., 12:15]
    xi = (rp - r - v - mu) / (np.sqrt(tau) * np.exp(s_f))[..., None]
    lqf = -0.5 * np.sum(xi * xi, -1) - 3 * s_f
    e = r - (rp + vr + mu_r)
    lqr = -np.sum(e * e, -1) / (2 * tau * np.exp(2 * s_r)) - 3 * s_r
    prob = tro[..., 0] ** 2 * np.exp(lqr - lqf)
    perr = np.max(np.abs(prob - lt[..., 10]) / np.maximum(prob, 1e-3))
    f = np.minimum(1.0, prob) * np.sum((rp - r) ** 2, -1)
    ferr = np.max(np.abs(f - lt[..., 11]) / np.maximum(np.abs(f), 1e-3))
    print(f"log q_fwd {np.max(np.abs(lqf - lt[..., 8])):.2e}, acceptance probability {perr:.2e}, objective {ferr:.2e}")
    assert np.max(np.abs(lqf - lt[..., 8])) < 1e-6 and np.max(np.abs(lqr - lt[..., 9])) < 1e-6
    assert perr < 1e-6 and ferr < 1e-6
    # decisions consistent with the probability: certain accepts / rejects
    assert np.all(tr[..., 1][lt[..., 10] >= 1.0] == 1)
    # the network output traced at r equals the reference network on the reference features
    Pd = unflat(c.mlp_params())
    for (a, b, i, w) in [(0, 0, 3, 1), (1, 0, 17, 4), (0, 0, 21, 0)]:
        Rw = replay_walker(R, tr, a, b, i, w)
        x = lp.features(s.L, s.ions, Rw, i, r[a, b, i, w], v[a, b, i, w], tau)
        out, _ = lp.forward(Pd, x[None].astype(np.float32).astype(np.float64), emulate=True)
        assert np.max(np.abs(out[0] - lt[a, b, i, w, 0:4])) < 2e-2 * max(1.0, np.max(np.abs(out)))


def relerr_ratio(tr, tro):
    return np.max(np.abs(tr[..., 0] - tro[..., 0]) / np.maximum(np.abs(tro[..., 0]), 1e-2))


def replay_walker(R0, tr, a, b, i, w):
    """positions of walker w just before traced move (a, b, i)"""
    R = R0[w].copy()
    steps, sub, N = tr.shape[:3]
    for aa in range(steps):
        for bb in range(sub):
            for ii in range(N):
                if (aa, bb, ii) == (a, b, i):
                    return R
                if tr[aa, bb, ii, w, 1] != 0:
                    R[ii] = tr[aa, bb, ii, w, 2:5]
    return R


def one_electron_system():
    """1 electron, 1 positive orbital (no nodes), J1 with 5 ions, cubic L = 6."""
    M, L = 8, 6.0
    g = np.arange(M) / M
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    coefs = (1.6 + np.cos(2 * np.pi * X) + 0.5 * np.sin(2 * np.pi * Y) + 0.3 * np.cos(2 * np.pi * Z))[..., None]
    rng = np.random.default_rng(3)
    p = jastrow_params(8)
    return System(L=np.array([L, L, L]), ions=rng.uniform(0, L, (5, 3)), n_up=1, n_dn=0, M=(M, M, M), norb=1,
                  coefs=coefs, j1_params=p * 0.7, j1_rcut=2.9, j1_cusp=0.0, j2_uu=p, j2_ud=p, j2_rcut=2.9,
                  j2_cusp_uu=-0.25, j2_cusp_ud=-0.5, delay_rank=1)


def test_exact_sampling_one_electron():
    """Metropolis-Hastings with the asymmetric correction samples |psi|^2 for any network: the x
    marginal of 1024 walkers x 250 steps against the exact marginal (oracle log psi on a 24^3
    grid), chi^2 per degree of freedom below 4 (chains are autocorrelated; a sampler that drops
    the correction is biased by the network's mean shift)."""
    s = one_electron_system()
    G = 24
    g = (np.arange(G) + 0.5) / G * s.L[0]
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 1, 3)
    o = orc.VMC(s, pts, 1, 1)
    lpsi, _, _ = o.evaluate_log()
    p = np.exp(2 * (lpsi - lpsi.max())).reshape(G, G, G).sum((1, 2))
    p /= p.sum()
    rng = np.random.default_rng(10)
    nw = 1024
    R0 = rng.uniform(0, s.L[0], (nw, 1, 3))
    c = ctx_for(s, R0)
    P = lp.init_params(rng, zero_head=False, scale=1.0)
    P["b4"][:] = [0.3, -0.2, 0.1, 0.2]  # a clear mean shift and a wider proposal
    c.mlp_params(flat(P))
    c.proposal_config(learned=True, train=False)
    c.vmc_run(50, 1, 0.5, 0, 11, energies=False)
    xs = []
    for _ in range(25):
        c.vmc_run(10, 1, 0.5, 0, 12, energies=False)
        xs.append(np.mod(c.walkers_get()[:, 0, 0], s.L[0]))
    h = np.histogram(np.concatenate(xs), bins=G, range=(0, s.L[0]))[0]
    expct = p * h.sum()
    chi2 = float(np.sum((h - expct) ** 2 / expct)) / (G - 1)
    print("chi^2 / dof of the x marginal:", chi2, "acceptance", c.proposal_stats())
    assert chi2 < 4.0


@pytest.mark.parametrize("precision", ["fp64", "fp32", "t32", "tbf16"])
def test_training_improves_objective(precision):
    """Training (Adam on -mean acceptance-weighted squared jump) for 20 sweeps of 256 walkers of the
    22-electron system; the trained network's objective exceeds the untrained (plain) proposal's, in
    every storage mode of the table and engine."""
    s = small_system()
    R = np.random.default_rng(13).uniform(0, 1, size=(256, s.N, 3)) * s.L
    c = ctx_for(s, R, precision)
    c.proposal_config(learned=True, train=False, init_seed=5)
    c.vmc_run(3, 1, 0.3, 0, 21, energies=False)
    c.proposal_stats()
    c.vmc_run(10, 1, 0.3, 0, 22, energies=False)
    plain = c.proposal_stats()
    c.proposal_config(learned=True, train=True, lr=2e-3)
    c.vmc_run(20, 1, 0.3, 0, 23, energies=False)
    c.proposal_stats()
    c.proposal_config(learned=True, train=False)
    c.vmc_run(10, 1, 0.3, 0, 24, energies=False)
    trained = c.proposal_stats()
    print("plain", plain, "trained", trained)
    assert trained["train_steps"] > 0
    assert trained["objective"] > 1.02 * plain["objective"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_learned_move_matches_five_launch_form(precision, monkeypatch):
    """The frozen-network learned move in one kernel (network evaluated per walker on row 0 of the
    matrix-core tile) reproduces the five-launch form (per-walker kernels around the batched
    network).  FP64: identical decisions and acceptance counts; the two kernels are compiled
    separately, so floating-point contraction may differ between them: energies within 1e-12
    relative, positions and traces within 1e-11 (measured 6.6e-15, 6.9e-15 and 2.4e-13).  With the FP32
    determinant engine the fused kernel keeps row d of Binv from the inverse-row update in FP64
    while the five-launch form reloads it from the FP32-stored matrix, so there the decisions are
    identical and the values agree at the FP32 engine's accuracy on these random orbitals (ratios
    5e-3 relative above 1e-2, positions 1e-3, energies 1e-4; measured 3.9e-4 for ratios)."""
    rng = np.random.default_rng(19)
    s = small_system(k=4)
    R = rng.uniform(0, 1, size=(5, s.N, 3)) * s.L
    P = flat(lp.init_params(rng, zero_head=False, scale=0.5))
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("QMI_LEARN_FUSED", fused)
        c = ctx_for(s, R, precision)
        c.mlp_params(P)
        c.proposal_config(learned=True, train=False)
        E, tr, lt = c.vmc_run_learned(2, 2, 0.4, 2, 55, trace=True)
        out.append((E, tr, lt, c.walkers_get(), c.vmc_stats()))
    if precision == "fp64":
        (E1, tr1, lt1, R1, _), (E0, tr0, lt0, R0, _) = out
        assert np.array_equal(tr1[..., 1], tr0[..., 1])
        assert np.max(np.abs(E1 - E0) / np.abs(E0)) < 1e-12
        assert np.max(np.abs(R1 - R0)) < 1e-11
        assert np.max(np.abs(tr1 - tr0)) < 1e-11 and np.max(np.abs(lt1 - lt0)) < 1e-11
    else:
        (E1, tr1, lt1, R1, _), (E0, tr0, lt0, R0, _) = out
        assert np.array_equal(tr1[..., 1], tr0[..., 1])
        err = dict(ratio=relerr_ratio(tr1, tr0), position=float(np.max(np.abs(R1 - R0))),
                   energy=float(np.max(np.abs(E1 - E0) / np.abs(E0))))
        print("fused vs five-launch (FP32 engine):", err)
        assert err["ratio"] < 5e-3 and err["position"] < 1e-3 and err["energy"] < 1e-4
    assert out[0][4] == out[1][4]
