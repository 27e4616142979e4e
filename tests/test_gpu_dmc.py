"""DMC (DMCBatched + SFNBranch + WalkerControl, fluctuating population) on the HIP path against the
CPU FP64 oracle's DMC driver (same walkers, same seeds).

The oracle's DMC is pinned by tests/test_oracle_dmc.py (fixed-node sign rule, the branch weight of
the first step, projection below the VMC energy, population feedback).  Both drivers draw the same
random numbers (walker slot w from seed + w, walker control from seed - 1), so the populations
must agree walker for walker: per step the walker counts and move counts exactly, the weighted
energy, its variance, E_trial, E_ref and the effective time step within 1e-10 relative; at the end
the positions (1e-10 absolute), weights and local energies (1e-10 relative).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import make_complex_system, make_system, make_walkers

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel(a, b, floor=1.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def dmc_pair(system, nw, steps, seed=21, move_kernel=0, **kw):
    R = make_walkers(system, nw, 5)
    zion = system.N / len(system.ions)
    c = qmi.Context(0)
    c.system_create(system)
    c.walkers_load(R, 1)
    c.tune_set(move_kernel=move_kernel)
    c.coulomb_setup(zion)
    c.dmc_init(seed=seed, **kw)
    st = c.dmc_run(steps)
    o = orc.VMC(system, R, 1, seed)
    o.dmc_init(zion, seed=seed, **kw)
    sto = o.dmc_run(steps)
    return c, o, st, sto


def check_dmc(c, o, st, sto):
    print("walkers per step", st[:, 4].astype(int).tolist())
    assert np.array_equal(st[:, [3, 4, 8, 9, 10]], sto[:, [3, 4, 8, 9, 10]])
    for col in (0, 2, 5, 6, 7):
        assert rel(st[:, col], sto[:, col]) < REL, col
    assert rel(st[:, 1], sto[:, 1], floor=max(1.0, float(np.max(sto[:, 1])) * 1e-3)) < 1e-9
    assert c.nw == o.nw
    assert np.allclose(c.walkers_get(), o.get()["R"], rtol=0, atol=1e-10)
    w, e = c.dmc_walker_data()
    wo, eo = o.dmc_walker_data()
    assert rel(w, wo) < REL and rel(e, eo) < REL


@pytest.mark.parametrize("k,warmup,history", [(1, 4, 0), (4, 0, 1), (6, 3, 0)])
def test_dmc_matches_oracle(k, warmup, history):
    s = make_system(L=7.0, n_ion=3, n_up=7, n_dn=6, norb=8, M=12, delay_rank=k, seed=13)
    c, o, st, sto = dmc_pair(s, 24, 12, tau=0.02, capacity=96, warmup_steps=warmup, refenergy_history=history,
                             recompute_period=5)
    check_dmc(c, o, st, sto)
    assert st[-1, 4] != st[0, 3] or np.any(st[:, 3] != st[:, 4])  # the population changed


def test_dmc_complex_orbitals():
    s = make_complex_system(L=7.0, n_ion=3, n_up=6, n_dn=5, norb=8, M=12, delay_rank=2, seed=17)
    c, o, st, sto = dmc_pair(s, 16, 10, tau=0.02, capacity=64, warmup_steps=2, cutoff_scheme=1)
    check_dmc(c, o, st, sto)


def test_dmc_production_shaped_system():
    """128 electrons (64 per spin), delay rank 16 (flushes within the sweep), 64 walkers."""
    s = make_system(L=10.0, n_ion=16, n_up=64, n_dn=64, norb=64, M=16, delay_rank=16, seed=19)
    c, o, st, sto = dmc_pair(s, 64, 6, tau=0.005, capacity=256, warmup_steps=2)
    check_dmc(c, o, st, sto)


def test_dmc_capacity_exceeded_is_an_error():
    s = make_system(L=7.0, n_ion=3, n_up=4, n_dn=4, norb=4, M=12, delay_rank=1, seed=23)
    R = make_walkers(s, 8, 2)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    with pytest.raises(RuntimeError, match="qmi_coulomb_setup"):
        c.dmc_init(tau=0.02, capacity=8, seed=3)
    c.coulomb_setup(2.0)
    # a target 100 times the population raises E_trial by log(100) / tau: the weights reach ~100
    c.dmc_init(tau=0.02, capacity=16, seed=3, warmup_steps=5, target_walkers=800)
    with pytest.raises(RuntimeError, match="exceeds the capacity"):
        c.dmc_run(5)
    with pytest.raises(RuntimeError, match="qmi_nlpp_setup"):
        c.dmc_init(tau=0.02, capacity=16, seed=3, nonlocal_moves=1)


def test_dmc_with_pseudopotential():
    """The local energy with the non-local pseudopotential term, each walker's quadrature grid
    rotated by three draws from its stream after the sweep (identical on both sides)."""
    from test_oracle_nlpp import channels, icosahedron
    s = make_system(L=7.0, n_ion=4, n_up=7, n_dn=6, norb=8, M=12, delay_rank=4, seed=29)
    R = make_walkers(s, 20, 7)
    zion = s.N / len(s.ions)
    ls = [0, 1, 2]
    dr, v, v2, _ = channels(ls, seed=4)
    knots, w = icosahedron()
    kw = dict(tau=0.02, capacity=80, warmup_steps=3, recompute_period=4)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.coulomb_setup(zion)
    c.nlpp_setup(ls, dr, v, v2, 2.2, knots, w)
    c.dmc_init(seed=51, **kw)
    st = c.dmc_run(8)
    o = orc.VMC(s, R, 1, 51)
    o.dmc_set_nlpp(ls, dr, v, v2, 2.2, knots, w)
    o.dmc_init(zion, seed=51, **kw)
    sto = o.dmc_run(8)
    check_dmc(c, o, st, sto)


@pytest.mark.parametrize("k", [1, 4])
def test_dmc_with_tmoves(k):
    """T-moves v0 after the local energy: each walker draws one more number from its stream,
    selects a candidate among the quadrature points of its electron-ion pairs (NonLocalTOperator::
    selectMove) and moves that electron when the ratio is non-zero; the accepted-move counts must
    agree step by step with the oracle's, and the populations walker for walker."""
    from test_oracle_nlpp import channels, icosahedron
    s = make_system(L=7.0, n_ion=4, n_up=7, n_dn=6, norb=8, M=12, delay_rank=k, seed=29)
    R = make_walkers(s, 24, 7)
    zion = s.N / len(s.ions)
    ls = [0, 1, 2]
    dr, v, v2, _ = channels(ls, seed=4)
    knots, w = icosahedron()
    kw = dict(tau=0.05, capacity=96, warmup_steps=3, recompute_period=4, nonlocal_moves=1)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.coulomb_setup(zion)
    c.nlpp_setup(ls, dr, v, v2, 2.2, knots, w)
    c.dmc_init(seed=57, **kw)
    st = c.dmc_run(8)
    o = orc.VMC(s, R, 1, 57)
    o.dmc_set_nlpp(ls, dr, v, v2, 2.2, knots, w)
    o.dmc_init(zion, seed=57, **kw)
    sto = o.dmc_run(8)
    print("T-moves per step", st[:, 10].astype(int).tolist())
    check_dmc(c, o, st, sto)
    assert st[:, 10].sum() >= 5


def test_dmc_tmoves_production_shaped():
    """T-moves on 128 electrons (64 per spin), delay rank 16, 96 walkers: more electrons move per
    step than the delay rank, so the protocol moves flush the delayed block inside the loop."""
    from test_oracle_nlpp import channels, icosahedron
    s = make_system(L=10.0, n_ion=16, n_up=64, n_dn=64, norb=64, M=16, delay_rank=16, seed=19)
    R = make_walkers(s, 96, 5)
    zion = s.N / len(s.ions)
    ls = [0, 1]
    dr, v, v2, _ = channels(ls, seed=6)
    knots, w = icosahedron()
    kw = dict(tau=0.1, capacity=384, warmup_steps=2, nonlocal_moves=1)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    c.coulomb_setup(zion)
    c.nlpp_setup(ls, dr, v, v2, 2.0, knots, w)
    c.dmc_init(seed=73, **kw)
    st = c.dmc_run(3)
    o = orc.VMC(s, R, 1, 73)
    o.dmc_set_nlpp(ls, dr, v, v2, 2.0, knots, w)
    o.dmc_init(zion, seed=73, **kw)
    sto = o.dmc_run(3)
    print("T-moves per step", st[:, 10].astype(int).tolist())
    check_dmc(c, o, st, sto)
    assert st[:, 10].max() > 16


def test_dmc_mixed_precision_runs():
    """DMC on the FP32 determinant engine: the same driver, finite energies, a living population
    and a weighted energy close to the FP64 run's at the first step."""
    s = make_system(L=7.0, n_ion=3, n_up=7, n_dn=6, norb=8, M=12, delay_rank=4, seed=13)
    R = make_walkers(s, 24, 5)
    zion = s.N / len(s.ions)
    out = []
    for prec in ("fp64", "fp32"):
        c = qmi.Context(0, precision=prec)
        c.system_create(s)
        c.walkers_load(R, 1)
        c.coulomb_setup(zion)
        c.dmc_init(tau=0.02, capacity=96, seed=21, warmup_steps=4)
        out.append(c.dmc_run(6))
    assert np.all(np.isfinite(out[1])) and np.all(out[1][:, 4] > 0)
    assert abs(out[1][0, 0] - out[0][0, 0]) < 1e-3 * max(1.0, abs(out[0][0, 0]))


def test_dmc_tmoves_mixed_precision_runs():
    """T-moves on the FP32 determinant engine and FP32 table: the same driver, finite statistics,
    accepted non-local moves, the first step's energy close to the FP64 run's."""
    from test_oracle_nlpp import channels, icosahedron
    s = make_system(L=7.0, n_ion=4, n_up=7, n_dn=6, norb=8, M=12, delay_rank=4, seed=29)
    R = make_walkers(s, 32, 7)
    zion = s.N / len(s.ions)
    ls = [0, 1, 2]
    dr, v, v2, _ = channels(ls, seed=4)
    knots, w = icosahedron()
    out = []
    for prec in ("fp64", "fp32"):
        c = qmi.Context(0, precision=prec)
        c.system_create(s)
        c.walkers_load(R, 1)
        c.coulomb_setup(zion)
        c.nlpp_setup(ls, dr, v, v2, 2.2, knots, w)
        c.dmc_init(tau=0.1, capacity=128, seed=81, warmup_steps=3, nonlocal_moves=1)
        out.append(c.dmc_run(4))
    print("T-moves per step", out[1][:, 10].astype(int).tolist())
    assert np.all(np.isfinite(out[1])) and np.all(out[1][:, 4] > 0) and out[1][:, 10].sum() >= 2
    assert abs(out[1][0, 0] - out[0][0, 0]) < 1e-3 * max(1.0, abs(out[0][0, 0]))


@pytest.mark.parametrize("k", [1, 4])
def test_dmc_window_kernel_matches_oracle(k):
    """DMC with one launch per delay window (k_sweep_segment, the benchmark's FP64 form): node
    rejection, prob_eff and the tau-weighted sums inside the move loop, against the oracle."""
    s = make_system(L=7.0, n_ion=3, n_up=7, n_dn=6, norb=8, M=12, delay_rank=k, seed=13)
    c, o, st, sto = dmc_pair(s, 24, 10, move_kernel=1, tau=0.02, capacity=96, warmup_steps=3, recompute_period=5)
    assert c.tune_get()["move_kernel"] == 1
    check_dmc(c, o, st, sto)


def test_dmc_tmoves_pair_list_matches_per_electron_form(monkeypatch):
    """DMC with T-moves on 128 electrons and 32 ions (about 40 electron-ion pairs per walker): the
    pair-list evaluation of the pseudopotential energy and of each electron's scan weight (point
    weights stored per pair, then summed per electron in the per-electron form's order) against one
    workgroup per electron.  The per-point arithmetic is the same code in both kernels (the compiler
    may fuse multiply-adds differently: ~1e-15), so populations, branching counts and T-move counts
    agree exactly and energies and positions to 1e-10."""
    from test_oracle_nlpp import channels, icosahedron
    s = make_system(L=10.0, n_ion=32, n_up=64, n_dn=64, norb=64, M=16, delay_rank=8, seed=3)
    R = make_walkers(s, 16, 5)
    ls = [0, 1]
    dr, v, v2, _ = channels(ls, seed=4)
    knots, w = icosahedron()
    kw = dict(tau=0.02, capacity=48, warmup_steps=2, recompute_period=3, nonlocal_moves=1)
    out = []
    for form in ("pairs", "electron"):
        monkeypatch.setenv("QMI_NLPP_FORM", form)
        c = qmi.Context(0)
        c.system_create(s)
        c.walkers_load(R, 1)
        c.coulomb_setup(s.N / len(s.ions))
        c.nlpp_setup(ls, dr, v, v2, 1.5, knots, w)
        c.dmc_init(seed=61, **kw)
        out.append((c.dmc_run(5), c.walkers_get(), c))
    (st, Rp, cp), (ste, Re, ce) = out
    print("T-moves per step", st[:, 10].astype(int).tolist(), "walkers", st[:, 4].astype(int).tolist())
    assert st[:, 10].sum() >= 5
    assert np.array_equal(st[:, [3, 4, 8, 9, 10]], ste[:, [3, 4, 8, 9, 10]])
    for col in (0, 2, 5, 6, 7):
        assert rel(st[:, col], ste[:, col]) < REL, col
    assert np.allclose(Rp, Re, rtol=0, atol=1e-10)
