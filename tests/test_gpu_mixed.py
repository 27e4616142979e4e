"""Mixed-precision forms of the HIP path against the FP64 oracle (per-quantity errors).

The GPU runs VMC with a narrowed spline table (FP32, or BF16 storage) and either an FP64 or an FP32
determinant engine; the FP64 oracle replays the GPU's trajectory (positions and decisions) and
recomputes every quantity.  Two references are used:
  * the oracle on the table rounded exactly as the device stores it (FP32, or BF16 by round to
    nearest even of the FP32 value): the difference is the device's arithmetic and storage;
  * the oracle on the original FP64 table: adds the table quantization (about 3e-8 relative for
    FP32 and 2e-3 for BF16 coefficients).
Errors are relative to max(|x|, floor): ratios floor 1e-2, gradients / Laplacians / energies 1.

An FP32 determinant engine stores the inverse Slater matrix in FP32, so a ratio (a dot product of an
inverse row with the orbital values) carries an error of about cond(A) x 6e-8 times its
cancellation.  The synthetic random-coefficient orbitals have cond(A) = 4e2..1e4 at n = 128, where
that is ~1e-3 absolute near small ratios; the FP32-engine bounds are therefore checked on a
well-conditioned plane-wave system (cond up to 1e3 at n = 16), and on the random system only the
measured errors are printed with gross bounds.  The FP64-engine (table-only) modes keep FP64
parity against the same-table oracle.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import System, jastrow_params, make_system, make_walkers

pytestmark = pytest.mark.gpu


def relerr(a, b, floor=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


round_table = orc.round_table  # the coefficients as the device stores them, in FP64


def plane_wave_system(n_up, n_dn, M=16, L=10.0, seed=1, k=8):
    """Real plane waves cos / sin(2 pi m.x / L) of the lowest |m| on an M^3 grid."""
    norb = max(n_up, n_dn)
    cand = sorted({(a, b, c) for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4)} - {(0, 0, 0)},
                  key=lambda m: (m[0] ** 2 + m[1] ** 2 + m[2] ** 2, m))
    funcs = [("c", (0, 0, 0))]
    for m in cand:
        if tuple(-v for v in m) in [f[1] for f in funcs]:
            continue
        funcs += [("c", m), ("s", m)]
        if len(funcs) >= norb:
            break
    g = np.arange(M) / M
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    coefs = np.empty((M, M, M, norb))
    for o, (t, m) in enumerate(funcs[:norb]):
        ph = 2 * np.pi * (m[0] * X + m[1] * Y + m[2] * Z)
        coefs[..., o] = np.cos(ph) if t == "c" else np.sin(ph)
    rng = np.random.default_rng(seed)
    p = jastrow_params(8)
    return System(L=np.array([L, L, L]), ions=rng.uniform(0, L, (8, 3)), n_up=n_up, n_dn=n_dn, M=(M, M, M),
                  norb=norb, coefs=coefs, j1_params=p * 0.7, j1_rcut=L / 2, j1_cusp=0.0, j2_uu=p, j2_ud=p * 1.3,
                  j2_rcut=L / 2, j2_cusp_uu=-0.25, j2_cusp_ud=-0.5, delay_rank=k)


def run_and_compare(s, R, prec, steps=2, seed=81):
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.mw_evaluateLog()
    E, tr = c.vmc_run(steps, 1, 0.3, 0, seed, trace=True)
    lp, G, L, _ = c.mw_evaluateGL(False)
    res = {}
    for name, table in (("same", round_table(s.coefs, prec)), ("fp64", s.coefs)):
        o = orc.VMC(dataclasses.replace(s, coefs=table), R, 1, seed)
        o.evaluate_log()
        Eo, tro = o.replay(tr, 1, 0.3, 0)
        lpo, Go, Lo = o.evaluate_log()
        res[name] = dict(ratio=relerr(tr[..., 0], tro[..., 0], 1e-2), grad=relerr(tr[..., 5:8], tro[..., 5:8], 1.0),
                         energy=relerr(E, Eo, 1.0), logpsi=relerr(lp, lpo, 1.0), G=relerr(G, Go, 1.0),
                         L=relerr(L, Lo, 1.0), decisions=int(np.sum(tr[..., 1] != tro[..., 1])),
                         moves=int(tr[..., 1].size))
    return res


def slater_cond(s, R, nw):
    out = []
    for w in range(nw):
        for i0, n in ((0, s.n_up), (s.n_up, s.n_dn)):
            M = np.array([orc.spo_vgl(s.L, s.coefs, R[w, i0 + i], n)[0] for i in range(n)])
            out.append(np.linalg.cond(M))
    return max(out)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_fp32_engine_well_conditioned(prec):
    """FP32 determinant engine (FP32 inverse, delayed rows and MFMA flush) on plane-wave orbitals,
    16 + 16 electrons, delay rank 8, 4 walkers, 2 sweeps: per-move errors at the FP32 scale."""
    s = plane_wave_system(16, 16)
    R = make_walkers(s, 4, 5)
    res = run_and_compare(s, R, prec)
    print(f"{prec} (cond {slater_cond(s, R, 4):.0f}): {res}")
    e = res["same"]
    assert e["decisions"] <= 1
    assert e["ratio"] < 1e-4 and e["grad"] < 1e-4 and e["energy"] < 1e-5
    # log psi accumulates the log of every accepted FP32-engine ratio over the run
    assert e["logpsi"] < 1e-5 and e["G"] < 1e-4 and e["L"] < 1e-4


@pytest.mark.parametrize("prec", ["t32", "tbf16"])
def test_table_only_modes_fp64_parity(prec):
    """FP32 / BF16 table with the FP64 engine, 128 + 128 electrons, 128 orbitals on a 16^3 table, delay
    rank 32: FP64 parity against the oracle on the same rounded table (ratios relative to
    max(|ratio|, 0.1) at 1e-9: cond(A) up to 1e4 here); the FP64-table differences are the table
    quantization."""
    s = make_system(L=14.0, n_ion=32, n_up=128, n_dn=128, norb=128, M=16, delay_rank=32, seed=41)
    R = make_walkers(s, 4, 42)
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.mw_evaluateLog()
    E, tr = c.vmc_run(2, 1, 0.3, 0, 81, trace=True)
    o = orc.VMC(dataclasses.replace(s, coefs=round_table(s.coefs, prec)), R, 1, 81)
    o.evaluate_log()
    Eo, tro = o.replay(tr, 1, 0.3, 0)
    same = dict(ratio=relerr(tr[..., 0], tro[..., 0], 0.1), grad=relerr(tr[..., 5:8], tro[..., 5:8], 1.0),
                energy=relerr(E, Eo, 1.0), decisions=int(np.sum(tr[..., 1] != tro[..., 1])))
    full = run_and_compare(s, R, prec)["fp64"]
    print(f"{prec}: vs same-table oracle {same}; vs FP64-table oracle {full}")
    assert same["decisions"] == 0
    assert same["ratio"] < 1e-9 and same["grad"] < 1e-9 and same["energy"] < 1e-9
    # vs the FP64 table: the FP32 table is close; BF16 storage (2^-9 relative per coefficient) changes
    # these random orbitals visibly (ratios near nodes by O(1) relative to 1e-2), so only its energy
    # is bounded
    if prec == "t32":
        assert full["ratio"] < 1e-4 and full["energy"] < 1e-6
    else:
        assert full["energy"] < 1e-2


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_fp32_engine_on_random_orbitals(prec):
    """The FP32 engine on the random-coefficient system (cond(A) up to 1e4): errors printed with the
    measured condition number; only gross bounds are asserted."""
    s = make_system(L=14.0, n_ion=32, n_up=128, n_dn=128, norb=128, M=16, delay_rank=32, seed=41)
    R = make_walkers(s, 4, 42)
    res = run_and_compare(s, R, prec)
    print(f"{prec} (cond {slater_cond(s, R, 4):.0f}): {res}")
    e = res["same"]
    assert e["decisions"] <= e["moves"] // 100
    assert e["energy"] < 1e-2 and e["logpsi"] < 1e-4
