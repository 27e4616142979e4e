"""Mixed-precision forms of the HIP path against the FP64 oracle (per-quantity errors).

The GPU runs VMC with FP32 (or BF16-stored) spline tables and an FP32 determinant engine; the FP64
oracle replays the GPU's trajectory (positions and decisions) and recomputes every quantity.
Two references are used:
  * the oracle on the table rounded exactly as the device stores it (FP32, or BF16 by round to
    nearest even of the FP32 value): the difference is the device's FP32 arithmetic and storage,
    bounded by the per-move target 1e-5 relative for ratios (relative to max(|ratio|, 1e-2));
  * the oracle on the original FP64 table: adds the table quantization (about 3e-8 relative for
    FP32 and 4e-3 for BF16 coefficients), reported and bounded loosely.
Decisions may differ from the FP64 oracle only at near-ties of u and the acceptance probability.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import make_system, make_walkers

pytestmark = pytest.mark.gpu

RATIO_FLOOR = 1e-2


def relerr(a, b, floor=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def round_table(coefs, prec):
    """The coefficients as the device stores them, returned in FP64."""
    f = np.asarray(coefs, dtype=np.float64).astype(np.float32)
    if prec == "fp32":
        return f.astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def replay_errors(s_ref, R, tr, E):
    o = orc.VMC(s_ref, R, 1, 81)
    o.evaluate_log()
    Eo, tro = o.replay(tr, 1, 0.3, 0)
    return dict(ratio=relerr(tr[..., 0], tro[..., 0], RATIO_FLOOR),
                grad=relerr(tr[..., 5:8], tro[..., 5:8], 1.0),
                energy=relerr(E, Eo, 1.0),
                decisions=int(np.sum(tr[..., 1] != tro[..., 1])), moves=int(tr[..., 1].size)), o


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mixed_precision_vs_oracle(prec):
    """128 + 128 electrons, 128 orbitals on a 16^3 table, delay rank 32, 4 walkers, 2 sweeps."""
    s = make_system(L=14.0, n_ion=32, n_up=128, n_dn=128, norb=128, M=16, delay_rank=32, seed=41)
    R = make_walkers(s, 4, 42)
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.walkers_load(R, 1)
    lp, G, L = c.mw_evaluateLog()
    E, tr = c.vmc_run(2, 1, 0.3, 0, 81, trace=True)
    s_same = dataclasses.replace(s, coefs=round_table(s.coefs, prec))
    same, o_same = replay_errors(s_same, R, tr, E)
    full, _ = replay_errors(s, R, tr, E)
    # gradients / Laplacians of the final configuration from the device's running state
    lp2, G2, L2, _ = c.mw_evaluateGL(False)
    lpo, Go, Lo = o_same.evaluate_log()
    gl = dict(logpsi=relerr(lp2, lpo, 1.0), G=relerr(G2, Go, 1.0), L=relerr(L2, Lo, 1.0))
    print(f"{prec}: vs same-table oracle {same}; GL {gl}; vs FP64-table oracle {full}")
    assert same["decisions"] <= max(1, same["moves"] // 1000)
    assert same["ratio"] < 1e-5 and same["grad"] < 1e-5 and same["energy"] < 1e-5
    assert gl["G"] < 1e-5 and gl["L"] < 1e-5 and gl["logpsi"] < 1e-6
    bound = 1e-4 if prec == "fp32" else 5e-2
    assert full["ratio"] < bound and full["grad"] < bound and full["energy"] < bound
