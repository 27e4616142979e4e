"""Optimized long-range breakup of the Coulomb potential (the reference's default "opt_breakup",
LRCoulombSingleton ESLER -> LRHandlerTemp<CoulombFunctor, LPQHIBasis>), restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

The potential 1/r is split as V_l(r) + V_s(r): V_l(r) = sum_n t_n h_n(r) for r < r_c and 1/r
beyond, on the locally piecewise-quintic Hermite basis h_n (15 knots on [0, r_c], three functions
per knot: value, slope, curvature); the t_n minimise the weighted misfit of V_l's Fourier transform
above k_c, so that the k-space sum can stop at k_c.

Functions and the reference code they follow:
  lpqhi_h, lpqhi_hintr2, lpqhi_c   Particle/LongRange/LPQHIBasis.h:65-86, LPQHIBasis.cpp:76-225
  fit_kpoints                      Particle/LongRange/LRBreakup.h:316-386 (SetupKVecs, AddKToList)
  breakup                          LRHandlerTemp.h:195-227 (InitBreakup), LRBreakup.h:71-186
                                   (DoBreakup: normal equations accumulated k by k in list order,
                                   SVD with singular values below eps * S_max dropped)
  fk, vlr_r0, vsr_k0, rvsr         LRHandlerTemp.h:98-108, 153-167, 172-178 (evaluate,
                                   evaluateSR_k0, evaluateLR_r0, evalFk), LRCoulombSingleton.cpp
                                   :32-56 (CoulombFunctor: Fk, Xk, integrate_r2); F_k below
                                   k delta = 2.5 by quadrature (lpqhi_c_quadrature), where the
                                   reference's recursion loses up to 1e-9

Pins (tests/test_oracle_breakup.py): the reference's known answers for CoulombPBCAA with this
breakup (test_coulomb_pbcAA.cpp:52-163: constants and energies of one charge in a unit cube and two
in a BCC cell), the Madelung constant, the basis' Hermite interpolation conditions, the Fourier
coefficients and r^2 integrals against adaptive quadrature, and independence of the total energy
from the k-space cutoff.
"""
from __future__ import annotations

import numpy as np

# value / slope / curvature shape polynomials in x = |r - r_i| / delta (LPQHIBasis.h:149-166); the
# left half of each function is the mirror image, odd for the slope function
SHAPE = np.array([[1.0, 0.0, 0.0, -10.0, 15.0, -6.0],
                  [0.0, 1.0, 0.0, -6.0, 8.0, -3.0],
                  [0.0, 0.0, 0.5, -1.5, 1.5, -0.5]])
MIRROR = np.array([1.0, -1.0, 1.0])
NUM_KNOTS = 15  # LRHandlerTemp.h:197


class Basis:
    """LPQHI basis on [0, rc] with nknots knots (3 * nknots functions, n = 3 * knot + kind)."""

    def __init__(self, L, rc, nknots=NUM_KNOTS):
        self.L = np.asarray(L, float)
        self.volume = float(np.prod(self.L))
        self.rc = float(rc)
        self.nknots = nknots
        self.delta = self.rc / (nknots - 1)
        self.size = 3 * nknots


def lpqhi_h(b: Basis, n, r):
    """h_n(r) (LPQHIBasis.h:65-86), vectorised over r."""
    r = np.asarray(r, float)
    i, a = divmod(n, 3)
    ra, rb, rc = b.delta * (i - 1), b.delta * i, min(b.rc, b.delta * (i + 1))
    left = (rb - r) / b.delta
    right = (r - rb) / b.delta
    poly = lambda x: np.polynomial.polynomial.polyval(x, SHAPE[a])
    out = np.where(r <= rb, MIRROR[a] * poly(left), poly(right))
    return np.where((r < ra) | (r > rc), 0.0, out)


def lpqhi_hintr2(b: Basis, n):
    """int_0^rc r^2 h_n(r) dr in closed form (LPQHIBasis.cpp:76-137)."""
    j, a = divmod(n, 3)
    d3 = b.delta ** 3
    S = SHAPE[a]
    if j == 0:
        return d3 * sum(S[i] / (i + 3) for i in range(6))
    if j == b.nknots - 1:
        p1, p2, p3, s = 1.0 / 3.0, -float(j), float(j * j), 0.0
        for i in range(6):
            s += S[i] * (p1 + p2 + p3)
            p1 *= (i + 3.0) / (i + 4.0)
            p2 *= (i + 2.0) / (i + 3.0)
            p3 *= (i + 1.0) / (i + 2.0)
        return s * d3 * MIRROR[a]
    if a != 1:
        p1, p2, s = 1.0 / 3.0, float(j * j), 0.0
        for i in range(6):
            s += S[i] * (p1 + p2)
            p1 *= (i + 3.0) / (i + 4.0)
            p2 *= (i + 1.0) / (i + 2.0)
        return 2.0 * d3 * s
    return 4.0 * j * d3 * sum(S[i] / (i + 2.0) for i in range(6))


def lpqhi_c(b: Basis, k):
    """c[n, k] = 4 pi / (k V) int_0^rc r sin(k r) h_n(r) dr for every basis function, vectorised
    over k (LPQHIBasis.cpp:140-225: the integrals of x^m e^{ikr} over each knot interval by the
    integration-by-parts recursions E+ / E-, combined into D+ / D-)."""
    k = np.asarray(k, float)
    d, nk = b.delta, b.nknots
    pref = 4.0 * np.pi / (k * b.volume)
    out = np.zeros((b.size, k.size))
    for i in range(nk):
        # E+_m = int_{r_i}^{r_i + delta} x^m e^{ikr} dr, E-_m the same over [r_i - delta, r_i]
        # with x = |r - r_i| / delta
        Ep = [-(1j / k) * (np.exp(1j * k * d) - 1.0) * np.exp(1j * k * d * i)]
        Em = [-(1j / k) * (np.exp(-1j * k * d) - 1.0) * np.exp(1j * k * d * i)]
        for m in range(1, 7):
            Ep.append(-(1j / k) * (np.exp(1j * k * d * (i + 1)) - m / d * Ep[-1]))
            sg = -1.0 if m & 1 else 1.0
            Em.append(-(1j / k) * (sg * np.exp(1j * k * d * (i - 1)) - m / d * Em[-1]))
        Dp = [pref * (d * Ep[m + 1].imag + i * d * Ep[m].imag) for m in range(6)]
        Dm = [-pref * (d * Em[m + 1].imag + i * d * Em[m].imag) for m in range(6)]
        for a in range(3):
            s = np.zeros(k.size)
            for m in range(6):
                sg = -1.0 if (a + m) & 1 else 1.0
                if i == 0:
                    s += SHAPE[a, m] * Dp[m]
                elif i == nk - 1:
                    s += SHAPE[a, m] * Dm[m] * sg
                else:
                    s += SHAPE[a, m] * (Dp[m] + Dm[m] * sg)
            out[3 * i + a] = s
    return out


_GL_X, _GL_W = np.polynomial.legendre.leggauss(40)


def lpqhi_c_quadrature(b: Basis, k):
    """c[n, k] by 40-point Gauss-Legendre on each polynomial piece, for k delta below ~2.5 (the
    long-range sum's k <= k_c): there each step of the E+ / E- recursion divides by k delta and the
    first shells lose up to 1e-9 relative; the integrand r sin(k r) h_n(r) is analytic on a piece
    spanning under 2.5 radians, where 40 nodes are exact to rounding."""
    k = np.atleast_1d(np.asarray(k, float))
    out = np.zeros((b.size, k.size))
    for n in range(b.size):
        i = n // 3
        for lo, hi in ((max(0.0, b.delta * (i - 1)), b.delta * i), (b.delta * i, min(b.rc, b.delta * (i + 1)))):
            if hi <= lo:
                continue
            r = 0.5 * (hi - lo) * _GL_X + 0.5 * (hi + lo)
            f = r * lpqhi_h(b, n, r) * (0.5 * (hi - lo) * _GL_W)
            out[n] += np.sin(np.outer(k, r)) @ f
    return out * (4.0 * np.pi / (k * b.volume))


def lattice_kpoints(L, kcut):
    """Every nonzero k = 2 pi m / L with |k| <= kcut, grouped into shells of equal |k|^2 (shell key
    floor(|k|^2 * 1e7), KContainer.cpp:110-117, 150-176, 226-265): (vectors [nk, 3] in shell order,
    |k|^2 of each shell, shell sizes)."""
    L = np.asarray(L, float)
    m = (np.floor(L * kcut / (2 * np.pi)) + 1).astype(int)
    axes = [2 * np.pi * np.arange(-mm, mm + 1) / l for mm, l in zip(m, L)]
    K = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    k2 = np.sum(K * K, 1)
    keep = (k2 > 0) & (k2 <= kcut * kcut)
    K, k2 = K[keep], k2[keep]
    key = np.floor(k2 * 1e7).astype(np.int64)
    order = np.argsort(key, kind="stable")
    K, k2, key = K[order], k2[order], key[order]
    first = np.concatenate([[True], key[1:] != key[:-1]])
    shell_k2 = k2[first]
    sizes = np.diff(np.concatenate([np.nonzero(first)[0], [k2.size]]))
    return K, shell_k2, sizes


def fit_kpoints(L, kc, kcont, kmax, nshells=4000):
    """The k points of the fit and their weights (LRBreakup.h:316-386): every lattice shell with
    |k|^2 above max(kc^2, first shell) up to kcont, weighted by its size; then nshells continuum
    shells of equal width up to kmax at their mid |k|, weighted by shell volume / (2 pi)^3 V^-1.
    Also returns the number of lattice shells at or below kc (the long-range sum's shells)."""
    V = float(np.prod(L))
    _, shell_k2, sizes = lattice_kpoints(L, kcont)
    kc2 = max(kc * kc, shell_k2[0])
    n_lr = int(np.sum(shell_k2 <= kc2))
    ks = list(np.sqrt(shell_k2[n_lr:]))
    ws = list(sizes[n_lr:].astype(float))
    kel = 8 * np.pi ** 3 / V
    dk = (kmax - kcont) / nshells
    for i in range(nshells):
        k1 = kcont + dk * i
        k2 = k1 + dk
        kk, w = 0.5 * (k1 + k2), 4.0 * np.pi * (k2 ** 3 - k1 ** 3) / 3.0 / kel
        # AddKToList merges a |k| equal to a listed one within 1e-12
        hit = np.nonzero(np.abs(np.asarray(ks) - kk) <= 1e-12)[0]
        if hit.size:
            ws[hit[0]] += w
        else:
            ks.append(kk)
            ws.append(w)
    return np.array(ks), np.array(ws), n_lr


class Breakup:
    """The fitted breakup of a cubic / orthorhombic cell (L = its sides)."""

    def __init__(self, L, dim_cutoff=15.0, nknots=NUM_KNOTS):
        L = np.asarray(L, float)
        self.rc = 0.5 * float(L.min())  # Wigner-Seitz radius of an orthorhombic cell
        self.kc = dim_cutoff / self.rc
        self.basis = b = Basis(L, self.rc, nknots)
        kcont = 60 * np.pi * b.volume ** (-1.0 / 3.0)  # LRHandlerTemp.h:205-206
        kmax = 6000.0 / self.rc
        k, w, self.n_lr_shells = fit_kpoints(L, self.kc, kcont, kmax)
        xk = -4 * np.pi / (b.volume * k * k) * np.cos(k * self.rc)  # CoulombFunctor::Xk
        cnk = lpqhi_c(b, k)
        # normal equations, accumulated one k point at a time in list order (the near-null space
        # of A, condition number ~1e11, follows this rounding)
        A = np.zeros((b.size, b.size))
        rhs = np.zeros(b.size)
        for q in range(k.size):
            c = cnk[:, q]
            rhs += w[q] * xk[q] * c
            A += w[q] * np.outer(c, c)
        U, s, Vt = np.linalg.svd(A)
        sinv = np.where(s < np.finfo(float).eps * s.max(), 0.0, 1.0 / np.where(s > 0, s, 1.0))
        self.t = Vt.T @ (sinv * (U.T @ rhs))
        self.chi2 = float(np.sum(w * (xk - self.t @ cnk) ** 2))
        self.nk_fit = k.size

    def fk(self, k):
        """F_k of the long-range sum: the transform of V_l (LRHandlerTemp.h:253-259 evalFk)."""
        k = np.atleast_1d(np.asarray(k, float))
        b = self.basis
        small = k * b.delta < 2.5
        c = np.empty((b.size, k.size))
        if np.any(small):
            c[:, small] = lpqhi_c_quadrature(b, k[small])
        if np.any(~small):
            c[:, ~small] = lpqhi_c(b, k[~small])
        return 4 * np.pi / (b.volume * k * k) * np.cos(k * self.rc) + self.t @ c

    def vlr_r0(self):
        return float(sum(self.t[n] * lpqhi_h(self.basis, n, 0.0) for n in range(self.basis.size)))

    def vsr_k0(self):
        b = self.basis
        v = 0.5 * self.rc ** 2 - sum(self.t[n] * lpqhi_hintr2(b, n) for n in range(b.size))
        return float(v * 4 * np.pi / b.volume)

    def rvsr(self, r):
        """r V_s(r) = 1 - r sum_n t_n h_n(r) below r_c, 0 beyond (LRHandlerTemp.h:98-108)."""
        r = np.asarray(r, float)
        vl = sum(self.t[n] * lpqhi_h(self.basis, n, r) for n in range(self.basis.size))
        return np.where(r >= self.rc, 0.0, 1.0 - r * vl)

    def lr_kvectors(self):
        """The k vectors of the long-range sum: the lattice shells at or below k_c."""
        K, shell_k2, sizes = lattice_kpoints(self.basis.L, self.kc)
        n = int(np.sum(sizes[: self.n_lr_shells]))
        return K[:n]

    def madelung(self):
        """Madelung constant of the cell: (sum_k F_k - v_l(0) - v_s(k=0)) / 2
        (CoulombPBCAA.cpp:621-624)."""
        K = self.lr_kvectors()
        return 0.5 * (float(np.sum(self.fk(np.sqrt(np.sum(K * K, 1))))) - self.vlr_r0() - self.vsr_k0())

    def spline(self, ng=1001):
        """r V_s(r) on the reference's linear grid of ng points over [0, r_c] as a clamped cubic
        spline (LRCoulombSingleton.cpp:137-158, CoulombPBCAA.cpp:467-476): first value
        extrapolated, last 0, end slopes from the first interval and 0."""
        from scipy.interpolate import CubicSpline
        x = np.linspace(0.0, self.rc, ng)
        y = self.rvsr(x)
        y[0] = 2 * y[1] - y[2]
        y[-1] = 0.0
        return CubicSpline(x, y, bc_type=((1, (y[1] - y[0]) / (x[1] - x[0])), (1, 0.0)))

    def vsr_fn(self, spline=False):
        """V_s(r) as the exact function or through the spline (zero at and beyond r_c)."""
        if not spline:
            return lambda r: self.rvsr(r) / r
        sp = self.spline()
        return lambda r: np.where(r >= self.rc, 0.0, sp(np.minimum(r, self.rc)) / r)


def _min_image(L, d):
    return d - L * np.round(d / L)


def coulomb_aa(L, r, q, dim_cutoff=15.0, brk: Breakup | None = None, spline=False):
    """CoulombPBCAA with this breakup: (total, (short range, long range, constants))
    (CoulombPBCAA.cpp:568-640); spline=True evaluates V_s through the 1001-point spline."""
    brk = brk or Breakup(L, dim_cutoff)
    L = np.asarray(L, float)
    r, q = np.asarray(r, float).reshape(-1, 3), np.asarray(q, float)
    vsr = brk.vsr_fn(spline)
    sr = 0.0
    for i in range(1, len(q)):
        d = _min_image(L, r[:i] - r[i])
        sr += q[i] * np.sum(q[:i] * vsr(np.sqrt(np.sum(d * d, 1))))
    K = brk.lr_kvectors()
    rho = np.exp(1j * (K @ r.T)) @ q
    lr = 0.5 * float(np.sum(brk.fk(np.sqrt(np.sum(K * K, 1))) * np.abs(rho) ** 2))
    consts = -0.5 * float(np.sum(q * q)) * brk.vlr_r0() - 0.5 * float(np.sum(q)) ** 2 * brk.vsr_k0()
    return sr + lr + consts, (sr, lr, consts)


def coulomb_ab(L, ra, qa, rb, qb, dim_cutoff=15.0, brk: Breakup | None = None, spline=False):
    """CoulombPBCAB with this breakup: (total, (short range, long range, constants)), constants
    -v_s(k=0) sum(qa) sum(qb) (CoulombPBCAB.cpp evalConsts)."""
    brk = brk or Breakup(L, dim_cutoff)
    L = np.asarray(L, float)
    ra, rb = np.asarray(ra, float).reshape(-1, 3), np.asarray(rb, float).reshape(-1, 3)
    qa, qb = np.asarray(qa, float), np.asarray(qb, float)
    vsr = brk.vsr_fn(spline)
    sr = 0.0
    for a in range(len(qa)):
        d = _min_image(L, rb - ra[a])
        sr += qa[a] * np.sum(qb * vsr(np.sqrt(np.sum(d * d, 1))))
    K = brk.lr_kvectors()
    rho_a = np.exp(1j * (K @ ra.T)) @ qa
    rho_b = np.exp(1j * (K @ rb.T)) @ qb
    lr = float(np.sum(brk.fk(np.sqrt(np.sum(K * K, 1))) * np.real(rho_a * np.conj(rho_b))))
    consts = -brk.vsr_k0() * float(np.sum(qa)) * float(np.sum(qb))
    return sr + lr + consts, (sr, lr, consts)
