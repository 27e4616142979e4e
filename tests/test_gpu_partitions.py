"""Walker partitions of the device-resident sweep (qmi_tune_partitions): the moves of P walker ranges
run on P streams forked from the context's stream.  Each walker's arithmetic is unchanged, so the
energies, positions, traces and acceptance counts must equal those of the unpartitioned sweep bit for
bit, eagerly and replayed from captured graphs, with delayed updates and flushes inside the sweep, for
complex orbitals, and in DMC (populations, weights and energies step by step).
"""
import numpy as np
import pytest

from qmcpack_amd import qmi
from qmcpack_amd.system import make_complex_system, make_system, make_walkers

pytestmark = pytest.mark.gpu


def run(s, R, P, graph, steps=3, sub=2, rp=2, trace=False, mk=0):
    c = qmi.Context(0)
    c.system_create(s)
    c.walkers_load(R, 1)
    assert c.tune_partitions(P) == P
    if mk:
        c.tune_set(move_kernel=mk)
    c.mw_evaluateLog()
    c.graph_enable(graph)
    E, tr = c.vmc_run(steps, sub, 0.35, rp, 91, trace=trace)
    return E, tr, c.walkers_get(), c.vmc_stats()


@pytest.mark.parametrize("k,P,graph", [(1, 2, False), (4, 3, False), (4, 3, True), (8, 4, True)])
def test_partitioned_sweep_is_bit_identical(k, P, graph):
    s = make_system(L=8.0, n_ion=4, n_up=11, n_dn=10, norb=12, M=12, delay_rank=k, seed=61)
    R = make_walkers(s, 64 * P + 5, 62)  # ragged last partition
    E1, tr1, R1, st1 = run(s, R, 1, graph, trace=not graph)
    EP, trP, RP, stP = run(s, R, P, graph, trace=not graph)
    assert np.array_equal(E1, EP)
    assert np.array_equal(R1, RP)
    assert st1 == stP
    if not graph:
        assert np.array_equal(tr1, trP)


@pytest.mark.parametrize("k,P", [(1, 2), (4, 3), (8, 2)])
def test_partitioned_window_kernel_is_bit_identical(k, P):
    """Partitions running the per-window kernel (one launch per delay window and partition, the
    flushes of each partition on its stream) against the unpartitioned per-window sweep."""
    s = make_system(L=8.0, n_ion=4, n_up=11, n_dn=10, norb=12, M=12, delay_rank=k, seed=61)
    R = make_walkers(s, 64 * P + 5, 62)
    E1, tr1, R1, st1 = run(s, R, 1, False, trace=True, mk=1)
    EP, trP, RP, stP = run(s, R, P, False, trace=True, mk=1)
    assert np.array_equal(E1, EP) and np.array_equal(R1, RP) and np.array_equal(tr1, trP)
    assert st1 == stP


def test_partitioned_complex_sweep_is_bit_identical():
    s = make_complex_system(L=7.0, n_ion=3, n_up=6, n_dn=5, norb=8, M=12, delay_rank=2, seed=63)
    R = make_walkers(s, 130, 64)
    E1, tr1, R1, _ = run(s, R, 1, False, trace=True)
    E2, tr2, R2, _ = run(s, R, 2, False, trace=True)
    assert np.array_equal(E1, E2) and np.array_equal(R1, R2) and np.array_equal(tr1, tr2)


@pytest.mark.parametrize("mk", [0, 1])
def test_partitioned_dmc_is_bit_identical(mk):
    s = make_system(L=7.0, n_ion=3, n_up=7, n_dn=6, norb=8, M=12, delay_rank=4, seed=65)
    R = make_walkers(s, 160, 66)
    out = []
    for P in (1, 2):
        c = qmi.Context(0)
        c.system_create(s)
        c.walkers_load(R, 1)
        c.tune_partitions(P)
        c.tune_set(move_kernel=mk)
        c.coulomb_setup(s.N / len(s.ions))
        c.dmc_init(tau=0.02, capacity=480, seed=17, warmup_steps=2)
        st = c.dmc_run(6)
        out.append((st, c.walkers_get(), c.dmc_walker_data()))
    (st1, R1, (w1, e1)), (st2, R2, (w2, e2)) = out
    assert np.array_equal(st1, st2)
    assert np.array_equal(R1, R2) and np.array_equal(w1, w2) and np.array_equal(e1, e2)


def test_partition_count_is_validated():
    c = qmi.Context(0)
    with pytest.raises(qmi.QmiError):
        c.tune_partitions(9)
    assert c.tune_partitions(0) == 1
