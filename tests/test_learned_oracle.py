"""Pins the CPU reference of the learned proposal (oracle/learned.py): features against a
brute-force definition, the MLP backward against central finite differences of the FP64 forward,
and the training objective's output gradients against finite differences on a synthetic log psi
with a known gradient."""
import numpy as np

from oracle import learned as lp


def test_features_nearest_sets():
    rng = np.random.default_rng(1)
    L = np.array([6.0, 7.0, 8.0])
    ions = rng.uniform(0, 1, (9, 3)) * L
    Rw = rng.uniform(0, 1, (25, 3)) * L
    p = Rw[3] + np.array([0.1, -0.2, 0.05])
    x = lp.features(L, ions, Rw, 3, p, np.array([0.1, 0.2, 0.3]), 0.3)
    # brute force: all minimum-image distances
    dI = [np.linalg.norm(lp.min_image(L, ions[j] - p)) for j in range(9)]
    dE = sorted(np.linalg.norm(lp.min_image(L, Rw[j] - p)) for j in range(25) if j != 3)
    assert np.allclose(x[3:16:4], sorted(dI)[:4])
    assert np.allclose(np.linalg.norm(x[16:64].reshape(16, 3), axis=1), dE[:16])
    assert np.allclose(x[64:], [0.1, 0.2, 0.3, 0.3])
    # translation of everything by a lattice vector leaves the features unchanged
    x2 = lp.features(L, ions + L * [1, -2, 0], Rw - L * [0, 1, 3], 3, p + L * [2, 0, 0], x[64:67], 0.3)
    assert np.allclose(x, x2)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    P = lp.init_params(rng, zero_head=False)
    X = rng.standard_normal((5, lp.N_IN))
    dout = rng.standard_normal((5, lp.N_OUT))
    _, cache = lp.forward(P, X, emulate=False)
    G = lp.backward(P, cache, dout)
    f = lambda: float(np.sum(lp.forward(P, X, emulate=False)[0] * dout))
    for name, idx in [("W1", (3, 60)), ("b1", (7,)), ("W2", (5, 9)), ("W3", (100, 127)), ("b3", (64,)),
                      ("W4", (2, 33)), ("b4", (3,))]:
        h = 1e-6
        P[name][idx] += h
        fp = f()
        P[name][idx] -= 2 * h
        fm = f()
        P[name][idx] += h
        fd = (fp - fm) / (2 * h)
        assert abs(fd - G[name][idx]) < 1e-6 * max(1.0, abs(fd)), (name, fd, G[name][idx])


def test_objective_output_grads_finite_differences():
    """Synthetic log psi(r') = a.r' - b |r'|^2 (gradient a - 2 b r'); ratio = psi(r')/psi(r)."""
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(3), 0.3
    logpsi = lambda x: float(a @ x - b * x @ x)
    grad = lambda x: a - 2 * b * x
    tau = 0.3
    r = rng.standard_normal(3)
    v = lp.unr_drift(tau, grad(r))
    xi = rng.standard_normal(3)
    v_rev = rng.standard_normal(3) * 0.2
    for trial in range(20):
        out_f = rng.standard_normal(4) * 0.3
        out_r = rng.standard_normal(4) * 0.3

        def f_of(of, orr):
            rp = r + v + of[:3] + np.sqrt(tau) * np.exp(of[3]) * xi
            ratio = np.exp(logpsi(rp) - logpsi(r))
            lq_f, lq_r = lp.log_q(tau, of, xi, r, rp, v_rev, orr)
            return min(1.0, ratio ** 2 * np.exp(lq_r - lq_f)) * float((rp - r) @ (rp - r))

        rp = r + v + out_f[:3] + np.sqrt(tau) * np.exp(out_f[3]) * xi
        ratio = np.exp(logpsi(rp) - logpsi(r))
        fval, gf, gr = lp.objective_output_grads(tau, ratio, grad(rp), xi, out_f, r, rp, v_rev, out_r)
        assert abs(fval - f_of(out_f, out_r)) < 1e-12 * max(1, fval)
        h = 1e-6
        for q in range(4):
            e = np.zeros(4)
            e[q] = h
            fd_f = (f_of(out_f + e, out_r) - f_of(out_f - e, out_r)) / (2 * h)
            fd_r = (f_of(out_f, out_r + e) - f_of(out_f, out_r - e)) / (2 * h)
            assert abs(fd_f - gf[q]) < 1e-5 * max(1.0, abs(fd_f)), (trial, q, fd_f, gf[q])
            assert abs(fd_r - gr[q]) < 1e-5 * max(1.0, abs(fd_r)), (trial, q, fd_r, gr[q])


def test_zero_head_is_plain_drift_diffusion():
    """With a zero output layer, log q_fwd / log q_rev equal the plain Green's functions times -1/(2 tau)
    normalisation (QMCDriverNew.h:268-297): log_gf = -|dr - v|^2/(2 tau), log_gb = -|r - r' - v'|^2/(2 tau)."""
    rng = np.random.default_rng(4)
    tau = 0.3
    r, v, v_rev, xi = (rng.standard_normal(3) for _ in range(4))
    out0 = np.zeros(4)
    rp = r + v + np.sqrt(tau) * xi
    lq_f, lq_r = lp.log_q(tau, out0, xi, r, rp, v_rev, out0)
    assert np.isclose(lq_f, -np.sum((rp - r - v) ** 2) / (2 * tau))
    assert np.isclose(lq_r, -np.sum((r - rp - v_rev) ** 2) / (2 * tau))
