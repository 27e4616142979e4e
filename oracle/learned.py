"""CPU reference of the learned per-electron proposal (test infrastructure only: imported by tests/
and by nothing on the product path).

The learned proposal has no counterpart in the reference.  Its definition (implemented on the
device by qmcpack_amd/csrc/qmi_mlp.hip and the learned-move kernels):

Features of electron i at position p in a walker (68 values, all in bohr):
  * the 4 nearest ions by minimum-image distance, ascending: (dx, dy, dz, |d|) with
    d = minimum image of (R_ion - p)  (the reference's source-minus-target convention,
    Particle/DistanceTable.h:365-371);
  * the 16 nearest other electrons, ascending: (dx, dy, dz) of minimum image (R_j - p);
  * the drift vector v at p (the UNR-scaled drift of DriftModifierUNR, a = 1);
  * the timestep tau.
Ties in distance are broken by the lower index.

Network: 68 -> 128 -> 128 -> 128 -> 4, SiLU hidden activations; the matrix products take BF16
operands (weights and layer inputs rounded to nearest even) with FP32 accumulation; bias and SiLU
in FP32.  Output (mu[3], s): proposal r' = p + v + mu + sigma xi, sigma = sqrt(tau) exp(s).
log q_fwd = -|xi|^2 / 2 - 3 s; the reverse density at r' uses the features and drift at r':
log q_rev = -|p - (r' + v' + mu')|^2 / (2 tau e^{2 s'}) - 3 s'   (shared constants cancel).
Acceptance min(1, |ratio|^2 exp(log q_rev - log q_fwd)).  With a zero output layer the proposal and
the acceptance are exactly the plain drift-diffusion ones (QMCDriverNew.h:268-297).

Training objective (per move): f = min(1, e^L) |r' - r|^2 with L = 2 ln|ratio| + log q_rev -
log q_fwd, maximised by gradient ascent with the reparameterisation r' = r + v + mu + sigma xi;
the features at r' and the drift v' are held fixed (no derivative through them), and
d(2 ln|psi(r')|)/dr' = 2 G(r') (the total gradient at the proposed position).
"""
import numpy as np

N_ION_NB, N_EL_NB = 4, 16
N_IN, N_HID, N_OUT = 68, 128, 4


def min_image(L, d):
    return d - L * np.round(d / L)


def features(L, ions, Rw, iat, p, drift, tau):
    """68 features of electron iat at position p in walker positions Rw [N][3]."""
    L = np.asarray(L, dtype=np.float64)
    x = np.zeros(N_IN)
    di = min_image(L, np.asarray(ions) - p)
    ri = np.sqrt(np.sum(di * di, axis=1))
    oi = np.lexsort((np.arange(len(ri)), ri))[:N_ION_NB]
    for k, j in enumerate(oi):
        x[4 * k:4 * k + 3] = di[j]
        x[4 * k + 3] = ri[j]
    de = min_image(L, np.asarray(Rw) - p)
    re = np.sqrt(np.sum(de * de, axis=1))
    idx = np.array([j for j in range(len(Rw)) if j != iat])
    oe = idx[np.lexsort((idx, re[idx]))][:N_EL_NB]
    for k, j in enumerate(oe):
        x[16 + 3 * k:16 + 3 * k + 3] = de[j]
    x[64:67] = drift
    x[67] = tau
    return x


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.asarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    sg = 1.0 / (1.0 + np.exp(-z))
    return sg * (1.0 + z * (1.0 - sg))


def init_params(rng, zero_head=True, scale=1.0):
    """W_l [out][in] (PyTorch-like uniform fan-in init), b_l [out]; a zero head reproduces plain
    drift-diffusion."""
    dims = [(N_HID, N_IN), (N_HID, N_HID), (N_HID, N_HID), (N_OUT, N_HID)]
    P = {}
    for l, (o, i) in enumerate(dims, 1):
        lim = scale / np.sqrt(i)
        P[f"W{l}"] = rng.uniform(-lim, lim, size=(o, i))
        P[f"b{l}"] = rng.uniform(-lim, lim, size=o)
    if zero_head:
        P["W4"][:] = 0.0
        P["b4"][:] = 0.0
    return P


def forward(P, X, emulate=True):
    """X [B][68] -> out [B][4] and the cache for backward.  emulate: BF16 operands of the matrix
    products, as on the device (accumulation in FP64 here: within FP32 rounding of the device)."""
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    h = np.asarray(X, dtype=np.float64)
    cache = {"h0": h}
    for l in (1, 2, 3):
        z = q(h) @ q(P[f"W{l}"]).T + P[f"b{l}"]
        if emulate:
            z = z.astype(np.float32).astype(np.float64)
        h = silu(z)
        if emulate:
            h = h.astype(np.float32).astype(np.float64)
        cache[f"z{l}"], cache[f"h{l}"] = z, h
    out = q(h) @ q(P["W4"]).T + P["b4"]
    return out, cache


def backward(P, cache, dout):
    """Gradients of sum_b dout[b] . out[b] with respect to the parameters (FP64; the device
    differentiates the same function in FP32 with the stored FP32 activations)."""
    G = {}
    d = np.asarray(dout, dtype=np.float64)
    G["W4"] = d.T @ cache["h3"]
    G["b4"] = d.sum(0)
    dh = d @ P["W4"]
    for l in (3, 2, 1):
        dz = dh * dsilu(cache[f"z{l}"])
        G[f"W{l}"] = dz.T @ cache[f"h{l - 1}"]
        G[f"b{l}"] = dz.sum(0)
        dh = dz @ P[f"W{l}"]
    return G


def unr_drift(tau, g):
    """DriftModifierUNR::getDrift with a = 1 (DriftModifierUNR.cpp)."""
    g = np.asarray(g, dtype=np.float64)
    v2 = float(g @ g)
    if v2 < np.finfo(float).eps:
        return tau * g
    sc = (-1.0 + np.sqrt(1.0 + 2.0 * tau * v2)) / v2
    return sc * g


def log_q(tau, out_f, xi, r, rp, v_rev, out_r):
    """log q_fwd and log q_rev (shared constants dropped)."""
    s_f = out_f[3]
    log_qf = -0.5 * float(xi @ xi) - 3.0 * s_f
    mu_r, s_r = out_r[:3], out_r[3]
    e = r - (rp + v_rev + mu_r)
    log_qr = -float(e @ e) / (2.0 * tau * np.exp(2.0 * s_r)) - 3.0 * s_r
    return log_qf, log_qr


def objective_output_grads(tau, ratio, G_new, xi, out_f, r, rp, v_rev, out_r):
    """f = min(1, e^L) |r' - r|^2 and df/d(out_f), df/d(out_r) (see the module docstring)."""
    s_f = out_f[3]
    sig_f = np.sqrt(tau) * np.exp(s_f)
    mu_r, s_r = out_r[:3], out_r[3]
    sig_r2 = tau * np.exp(2.0 * s_r)
    e = r - (rp + v_rev + mu_r)
    log_qf, log_qr = log_q(tau, out_f, xi, r, rp, v_rev, out_r)
    Lg = 2.0 * np.log(abs(ratio)) + log_qr - log_qf
    A = min(1.0, np.exp(Lg))
    D = rp - r
    D2 = float(D @ D)
    below = Lg < 0.0
    dL_dD = 2.0 * np.asarray(G_new) + e / sig_r2
    df_dD = 2.0 * A * D + (A * D2 * dL_dD if below else 0.0)
    gf = np.zeros(4)
    gr = np.zeros(4)
    gf[:3] = df_dD
    gf[3] = float(df_dD @ (sig_f * np.asarray(xi))) + (A * D2 * 3.0 if below else 0.0)
    if below:
        gr[:3] = A * D2 * e / sig_r2
        gr[3] = A * D2 * (float(e @ e) / sig_r2 - 3.0)
    return A * D2, gf, gr
