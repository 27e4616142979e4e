"""Uneven batching: captured-graph sub-steps, batches of contexts of different systems, tuning
knobs and the autotuner.

A graph replays the same kernels with the same arguments as the eagerly launched sub-step, and a
batch runs each context's kernels on its own stream; both must reproduce the eager single-context
run of every context bit for bit (energies, positions, acceptance counts).  The tuning knobs
(delay rank, move-kernel form, flush form) change only rounding: a tuned context is checked
against the oracle's replay of its trajectory like the untuned one (tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd import qmi
from qmcpack_amd.system import make_complex_system, make_system, make_walkers, uneven_populations

from test_gpu_parity import check_vmc

pytestmark = pytest.mark.gpu


def ctx_for(system, R, wpc=1):
    c = qmi.Context(0)
    c.system_create(system)
    c.walkers_load(R, wpc)
    c.mw_evaluateLog()
    return c


def small_pops():
    return uneven_populations(nw=24, sizes=(16, 32, 64), grids=(12, 14, 16), complex_fraction=0.25, seed=5,
                              delay_rank=8)


@pytest.mark.parametrize("kind,wpc", [("real", 1), ("real", 3), ("complex", 1)])
def test_graph_replay_matches_eager(kind, wpc):
    make = make_system if kind == "real" else make_complex_system
    s = make(L=9.0, n_ion=8, n_up=20, n_dn=18, norb=20, M=14, delay_rank=4, seed=11)
    R = make_walkers(s, 6, 12)
    out = []
    for graph in (False, True):
        c = ctx_for(s, R, wpc)
        c.graph_enable(graph)
        E = c.vmc_run(3, 2, 0.4, 2, 91)[0]
        out.append((E, c.walkers_get(), c.vmc_stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("concurrent", [True, False])
def test_batch_matches_individual_runs(concurrent):
    pops = small_pops()
    Rs = [make_walkers(s, n, 30 + i) for i, (s, n, _) in enumerate(pops)]
    seeds = [100 + 7 * i for i in range(len(pops))]
    alone = []
    for (s, n, _), R, sd in zip(pops, Rs, seeds):
        c = ctx_for(s, R)
        alone.append((c.vmc_run(2, 2, 0.3, 2, sd)[0], c.walkers_get(), c.vmc_stats()))
    ctxs = [ctx_for(s, R) for (s, n, _), R in zip(pops, Rs)]
    b = qmi.Batch(ctxs)
    b.set_concurrent(concurrent)
    E = b.vmc_run(2, 2, 0.3, 2, seeds)
    for i, c in enumerate(ctxs):
        assert np.array_equal(E[i], alone[i][0]), pops[i][2]
        assert np.array_equal(c.walkers_get(), alone[i][1]), pops[i][2]
        assert c.vmc_stats() == alone[i][2]


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_tuned_knobs_keep_parity(kind):
    """Delay rank 8 with the per-window move kernel and the three-pass flush (real), delay rank 16
    (complex): the oracle replays the trajectory within the FP64 tolerances."""
    make = make_system if kind == "real" else make_complex_system
    s = make(L=9.0, n_ion=8, n_up=20, n_dn=18, norb=20, M=14, delay_rank=4, seed=13)
    R = make_walkers(s, 4, 14)
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    if kind == "real":
        c.tune_set(8, 1, 1)
        assert c.tune_get() == dict(delay_rank=8, move_kernel=1, flush_variant=1)
        k = 8
    else:
        c.tune_set(16, 0, 0)
        k = 16
    c.mw_evaluateLog()
    E, tr = c.vmc_run(2, 2, 0.4, 2, 92, trace=True)
    import dataclasses
    o = orc.VMC(dataclasses.replace(s, delay_rank=k), R, 1, 92)
    o.evaluate_log()
    Eo, tro = o.replay(tr, 2, 0.4, 2)
    check_vmc(c, o, E, tr, Eo, tro)


def test_autotune_small_uneven_batch():
    pops = small_pops()
    ctxs = [ctx_for(s, make_walkers(s, n, 40 + i)) for i, (s, n, _) in enumerate(pops)]
    b = qmi.Batch(ctxs)
    choice, ms = b.autotune(0.3, trials=2)
    print("autotune:", [(p[2], tuple(ch)) for p, ch in zip(pops, choice)], ms)
    for (s, n, label), ch, c in zip(pops, choice, ctxs):
        nmax = max(s.n_up, s.n_dn)
        assert ch[0] in (1, 8, 16, 32, 64) and (ch[0] == 1 or ch[0] <= nmax)
        assert c.tune_get() == dict(delay_rank=int(ch[0]), move_kernel=int(ch[1]), flush_variant=int(ch[2]))
        if c.cplx:
            assert ch[1] == 0 and ch[2] == 0
    assert np.all(np.isfinite(ms)) and np.all(ms > 0)
    E = b.vmc_run(2, 1, 0.3, 0, [1 + i for i in range(len(ctxs))])
    assert all(np.all(np.isfinite(e)) for e in E)


def test_batch_graph_rebuilt_after_context_changes():
    """A context retuned (buffers reallocated) and run alone between two batched runs: the batch
    must capture again rather than replay kernels bound to the freed buffers.  Reference: the same
    sequence with every sub-step launched eagerly, context by context."""
    pops = small_pops()[:3]
    Rs = [make_walkers(s, n, 50 + i) for i, (s, n, _) in enumerate(pops)]
    seeds = [7, 8, 9]

    def sequence(batched):
        ctxs = [ctx_for(s, R) for (s, n, _), R in zip(pops, Rs)]
        out = []
        if batched:
            b = qmi.Batch(ctxs)
            out += b.vmc_run(1, 1, 0.3, 0, seeds)
        else:
            out += [c.vmc_run(1, 1, 0.3, 0, sd)[0] for c, sd in zip(ctxs, seeds)]
        ctxs[0].tune_set(4, 0, 1)
        ctxs[0].graph_enable(batched)
        out.append(ctxs[0].vmc_run(1, 1, 0.3, 0, seeds[0])[0])
        if batched:
            out += b.vmc_run(1, 1, 0.3, 0, seeds)
        else:
            out += [c.vmc_run(1, 1, 0.3, 0, sd)[0] for c, sd in zip(ctxs, seeds)]
        return out, [c.walkers_get() for c in ctxs]

    (Eb, Rb), (Ee, Re) = sequence(True), sequence(False)
    for a, b in zip(Eb, Ee):
        assert np.array_equal(a, b)
    for a, b in zip(Rb, Re):
        assert np.array_equal(a, b)


@pytest.mark.timeout(400)
def test_batch_at_workload_sizes_matches_oracle():
    """The batched run itself against the oracle at the uneven workload's sizes: 128 / 256 / 512 /
    1024 electrons on 48 / 64 / 64 / 80^3 tables, a real and a complex system each, delay rank 32,
    one walker per system, two sweeps with a from-scratch recompute after the second.

    The oracle replays the trajectory move by move (the chain is chaotic, so two FP64
    implementations cannot be compared free-running).  The per-move record it follows comes from
    each context run again alone, traced, from the same start and seed; that run must reproduce the
    batch's energies and positions bit for bit.  Against the oracle: identical accept/reject
    decisions, final positions and acceptance counts; energies within 1e-10 (relative to
    max(|E|, 1)); ratios held to the FP64 oracle's own distance from the long-double determinant
    shadow (within twice it, or 1e-10 if larger; relative to max(|ratio|, 1e-2)), and from the
    oracle by no more than the two distances from that shadow (plus 10 % and 1e-10): at 256
    electrons both FP64 implementations are 1e-9 from it, the device the closer one."""
    from concurrent.futures import ThreadPoolExecutor

    from test_gpu_parity import REL, relerr
    pops = uneven_populations(nw=8, sizes=(128, 256, 512, 1024), grids=(48, 64, 64, 80), complex_fraction=0.5,
                              seed=31, delay_rank=32)
    assert [n for _, n, _ in pops] == [1] * 8
    Rs = [make_walkers(s, n, 60 + i) for i, (s, n, _) in enumerate(pops)]
    seeds = [300 + 11 * i for i in range(len(pops))]
    steps, sub, tau, period = 2, 1, 0.3, 2
    ctxs = [ctx_for(s, R) for (s, n, _), R in zip(pops, Rs)]
    b = qmi.Batch(ctxs)
    Eb = b.vmc_run(steps, sub, tau, period, seeds)
    Rb = [c.walkers_get() for c in ctxs]
    Sb = [c.vmc_stats() for c in ctxs]
    b.close()
    traces = []
    for i, ((s, n, label), R, sd, c) in enumerate(zip(pops, Rs, seeds, ctxs)):
        c.walkers_load(R, 1)
        c.mw_evaluateLog()
        E, tr = c.vmc_run(steps, sub, tau, period, sd, trace=True)
        assert np.array_equal(E, Eb[i]) and np.array_equal(c.walkers_get(), Rb[i]), label
        traces.append(tr)
        c.close()

    def replay(i):
        s, R, sd = pops[i][0], Rs[i], seeds[i]
        o = orc.VMC(s, R, 1, sd)
        o.evaluate_log()
        Eo, tro, ex = o.replay(traces[i], sub, tau, period, exact=True)
        return o.get(), Eo, tro, ex

    with ThreadPoolExecutor(len(pops)) as pool:
        results = list(pool.map(replay, range(len(pops))))
    for i, ((s, n, label), (og, Eo, tro, ex)) in enumerate(zip(pops, results)):
        tr = traces[i]
        err = dict(decisions=int(np.sum(tr[..., 1] != tro[..., 1])), ratio=relerr(tr[..., 0], tro[..., 0], 1e-2),
                   ratio_exact=relerr(tr[..., 0], ex[..., 0], 1e-2),
                   oracle_exact=relerr(tro[..., 0], ex[..., 0], 1e-2),
                   grad=relerr(tr[..., 5:8], tro[..., 5:8], 1.0), energy=relerr(Eb[i], Eo, 1.0))
        print(label, err)
        assert err["decisions"] == 0, label
        assert err["ratio_exact"] <= max(REL, 2 * err["oracle_exact"]), label
        assert err["ratio"] <= 1.1 * (err["ratio_exact"] + err["oracle_exact"]) + REL, label
        assert err["grad"] < 1e-9 and err["energy"] < REL, label
        assert np.array_equal(Rb[i], og["R"]), label
        assert Sb[i]["accepted"] == og["accepted"] and Sb[i]["rejected"] == og["rejected"], label
