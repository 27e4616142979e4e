"""Restart files (walkers, random streams, DMC branch state) on the HIP path.

A run interrupted by checkpoint_save and continued in a fresh context after checkpoint_load is
checked against the CPU FP64 oracle: for VMC the oracle starts from the checkpoint's walkers and
streams and replays the restarted run's trace (every decision identical; ratios, gradients,
proposals and local energies within 1e-10 relative); for DMC (small time step, free-running) the
restarted steps follow the oracle's uninterrupted run: identical walker and acceptance counts,
positions within 1e-10 absolute, energies, weights, E_trial, E_ref and tau_eff within 1e-10
relative.  The restarted walkers are re-evaluated from scratch, as QMCDriverNew::
initialLogEvaluation does.  The saved random streams are pinned against numpy's MT19937 with the
legacy seeding (the C++ standard's mt19937 seeding and generation): a stream seeded with s after k
draws has the same 624 state words and position, and streams written by numpy drive a run.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import make_system, make_walkers

from test_gpu_parity import check_vmc

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel(a, b, floor=1.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def system(k=4, seed=13, n_ion=3):
    return make_system(L=7.0, n_ion=n_ion, n_up=7, n_dn=6, norb=8, M=12, delay_rank=k, seed=seed)


def vmc_ctx(s, R, wpc, prec="fp64"):
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.walkers_load(R, wpc)
    c.mw_evaluateLog()
    return c


def np_state(seed, draws=0):
    """numpy MT19937 (legacy seeding) after `draws` 32-bit outputs: (key[624], pos)."""
    bg = np.random.RandomState(seed)._bit_generator
    if draws:
        bg.random_raw(draws)
    st = bg.state["state"]
    return st["key"].astype(np.uint32), int(st["pos"])


@pytest.mark.parametrize("wpc", [1, 3])
def test_vmc_restart_replays_on_the_oracle(wpc):
    """Checkpoint after 3 sweeps, restart in a fresh context, 3 more sweeps with a trace; the oracle,
    started from the checkpoint's walkers and streams, replays that trace (the drift-diffusion chain
    amplifies rounding, so two FP64 implementations are compared on one trajectory): every
    decision, proposal, ratio, gradient and local energy as in the uninterrupted parity tests."""
    s = system()
    R = make_walkers(s, 6, 5)
    b = vmc_ctx(s, R, wpc)
    b.vmc_run(3, 2, 0.4, 3, seed=77)
    data = b.checkpoint_save()
    d = qmi.read_checkpoint(data)
    assert np.array_equal(d["state_0/walkers"], b.walkers_get())
    assert d["state_0/number_of_walkers"][0] == 6 and d["state_0/block"][0] == 3
    assert np.array_equal(d["state_0/walker_partition"], [0, 6])
    assert np.all(d["state_0/walker_weights"] == 1.0)
    assert d["random/std::mt19937"].shape == (-(-6 // wpc), 625)
    assert list(d["nprocs_nthreads_statesize"]) == [1, -(-6 // wpc), 625]

    c = qmi.Context(0)
    c.system_create(s)
    c.checkpoint_load(data)
    assert c.nw == 6 and np.array_equal(c.walkers_get(), b.walkers_get())
    # the checkpoint's seed continues the restored streams; the step count (3) keeps the
    # re-inversion schedule: after the third sweep here as after the oracle's third
    E, tr = c.vmc_run(3, 2, 0.4, 3, seed=77, trace=True)
    o = orc.VMC(s, d["state_0/walkers"], wpc, 1)
    o.set_rng_states(d["random/std::mt19937"])
    o.evaluate_log()
    Eo, tro = o.replay(tr, 2, 0.4, 3)
    check_vmc(c, o, E, tr, Eo, tro)


def test_saved_streams_are_mt19937_states():
    s = system(k=1)
    R = make_walkers(s, 4, 2)
    c = vmc_ctx(s, R, 1)
    c.vmc_run(0, seed=123)  # seeds the streams, draws nothing
    rs = qmi.read_checkpoint(c.checkpoint_save())["random/std::mt19937"]
    for w in range(4):
        key, pos = np_state(123 + w)
        assert np.array_equal(rs[w, :624], key) and rs[w, 624] == pos == 624
    c.vmc_run(1, seed=123)
    rs = qmi.read_checkpoint(c.checkpoint_save())["random/std::mt19937"]
    # one sweep: 3N Gaussians drawn in pairs, then one uniform per Metropolis test that reaches it
    k0 = 2 * ((3 * s.N + 1) // 2)
    for w in range(4):
        found = [k for k in range(k0, k0 + s.N + 1)
                 if np.array_equal(rs[w, :624], np_state(123 + w, k)[0]) and rs[w, 624] == np_state(123 + w, k)[1]]
        assert len(found) == 1, w


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_loaded_streams_drive_the_run(prec):
    """Streams written by numpy (fresh seeds 999 + w) replace the saved ones: the restarted run equals
    a run seeded with 999 on the same walkers, bit for bit (FP64 and the FP32 engine)."""
    s = system(k=2)
    R = make_walkers(s, 4, 3)
    b = vmc_ctx(s, R, 1, prec)
    b.vmc_run(2, seed=5)
    d = qmi.read_checkpoint(b.checkpoint_save())
    rs = np.zeros((4, 625), dtype=np.uint32)
    for w in range(4):
        rs[w, :624], rs[w, 624] = np_state(999 + w)
    d["random/std::mt19937"] = rs
    c = qmi.Context(0, precision=prec)
    c.system_create(s)
    c.checkpoint_load(qmi.write_checkpoint(d))
    E, _ = c.vmc_run(2, seed=5)  # the checkpoint's seed: continue the loaded streams
    ref = vmc_ctx(s, d["state_0/walkers"], 1, prec)
    Er, _ = ref.vmc_run(2, seed=999)
    assert np.array_equal(c.walkers_get(), ref.walkers_get()) and np.array_equal(E, Er)


@pytest.mark.parametrize("nlpp,warmup,history,tmoves", [(False, 3, 0, 0), (False, 7, 0, 0), (True, 2, 1, 0),
                                                        (True, 3, 0, 1)])
def test_dmc_restart_follows_the_uninterrupted_run(nlpp, warmup, history, tmoves):
    """Checkpoint after 5 steps (in the main stage with a limited energy history, during warmup,
    with the pseudopotential term and the unlimited history, or with T-moves), restart in a fresh
    context, 5 more steps against the oracle's 10 uninterrupted ones."""
    from test_oracle_nlpp import channels, icosahedron
    s = system(k=4, seed=29, n_ion=4)
    R = make_walkers(s, 20, 7)
    zion = s.N / len(s.ions)
    ls = [0, 1, 2]
    dr, v, v2, _ = channels(ls, seed=4)
    knots, w = icosahedron()
    kw = dict(tau=0.02, capacity=96, warmup_steps=warmup, refenergy_history=history, recompute_period=4,
              nonlocal_moves=tmoves)

    def ctx():
        c = qmi.Context(0)
        c.system_create(s)
        c.coulomb_setup(zion)
        if nlpp:
            c.nlpp_setup(ls, dr, v, v2, 2.2, knots, w)
        return c

    b = ctx()
    b.walkers_load(R, 1)
    b.dmc_init(seed=51, **kw)
    b.dmc_run(5)
    data = b.checkpoint_save()
    d = qmi.read_checkpoint(data)
    wb, eb = b.dmc_walker_data()
    assert np.array_equal(d["state_0/walker_weights"], wb) and np.array_equal(d["dmc/local_energies"], eb)
    assert d["random/std::mt19937"].shape[0] >= b.nw and d["random_master/std::mt19937"].shape == (1, 625)

    c = ctx()
    c.checkpoint_load(data)
    st = c.dmc_run(5)

    o = orc.VMC(s, R, 1, 51)
    if nlpp:
        o.dmc_set_nlpp(ls, dr, v, v2, 2.2, knots, w)
    o.dmc_init(zion, seed=51, **kw)
    o.dmc_run(5)
    sto = o.dmc_run(5)
    print("walkers per step", st[:, 4].astype(int).tolist())
    assert np.array_equal(st[:, [3, 4, 8, 9, 10]], sto[:, [3, 4, 8, 9, 10]])
    for col in (0, 2, 5, 6, 7):
        assert rel(st[:, col], sto[:, col]) < REL, col
    assert rel(st[:, 1], sto[:, 1], floor=max(1.0, float(np.max(sto[:, 1])) * 1e-3)) < 1e-9
    assert c.nw == o.nw
    assert np.allclose(c.walkers_get(), o.get()["R"], rtol=0, atol=1e-10)
    wc, ec = c.dmc_walker_data()
    wo, eo = o.dmc_walker_data()
    assert rel(wc, wo) < REL and rel(ec, eo) < REL


def test_checkpoint_errors_leave_the_context_unchanged():
    s = system(k=4)
    R = make_walkers(s, 5, 4)
    c = vmc_ctx(s, R, 1)
    c.vmc_run(1, seed=9)
    good = c.checkpoint_save()
    R0 = c.walkers_get()
    d = qmi.read_checkpoint(good)
    bad_pos = dict(d)
    bad_pos["random/std::mt19937"] = d["random/std::mt19937"].copy()
    bad_pos["random/std::mt19937"][0, 624] = 1000
    other = make_system(L=7.0, n_ion=3, n_up=5, n_dn=4, norb=6, M=12, delay_rank=1, seed=2)
    o = qmi.Context(0)
    o.system_create(other)
    o.walkers_load(make_walkers(other, 2, 1), 1)
    cases = [(b"not a checkpoint at all", "bad magic"), (good[:-9], "truncated"),
             (good + b"\0", "trailing bytes"), (qmi.write_checkpoint(bad_pos), "invalid random stream state")]
    for data, msg in cases:
        with pytest.raises(RuntimeError, match=msg):
            c.checkpoint_load(data)
    with pytest.raises(RuntimeError, match="do not have 9 electrons"):
        o.checkpoint_load(good)
    assert c.nw == 5 and np.array_equal(c.walkers_get(), R0)
    # a DMC checkpoint needs the Coulomb set-up (and the pseudopotential when it has the term)
    c.coulomb_setup(s.N / len(s.ions))
    c.dmc_init(tau=0.02, capacity=16, seed=3, warmup_steps=2)
    c.dmc_run(1)
    dmc = c.checkpoint_save()
    fresh = qmi.Context(0)
    fresh.system_create(s)
    with pytest.raises(RuntimeError, match="qmi_coulomb_setup"):
        fresh.checkpoint_load(dmc)
    # pending delayed updates (rank 4) cannot be saved
    p = vmc_ctx(s, R, 1)
    p.mw_evalGrad(0)
    p.mw_makeMove(0, np.full((5, 3), 0.05))
    p.mw_calcRatioGrad(0)
    p.mw_accept_rejectMove(0, np.ones(5, dtype=np.int32))
    with pytest.raises(RuntimeError, match="pending delayed updates"):
        p.checkpoint_save()
    p.mw_completeUpdates()
    assert qmi.read_checkpoint(p.checkpoint_save())["state_0/number_of_walkers"][0] == 5
