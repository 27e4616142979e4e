"""The oracle's non-local moves (T-moves v0: NonLocalECPotential::makeNonLocalMovesPbyP with
NonLocalTOperator::selectMove) pinned independently of its own arithmetic:

* selectMove against the reference's unit test (QMCHamiltonians/tests/test_NonLocalTOperator.cpp)
  and against a numpy formulation (cumulative weights, searchsorted) on random candidate lists;
* in DMC the move comes after the local energy: the first step's energies, weights, E_trial, E_ref
  and tau_eff equal those of the same run without T-moves (the first step does not branch);
* a T-move puts an electron at r_I + r u_j, on the sphere of its own distance r < r_cut to ion I:
  after the first step each walker differs from the run without T-moves in at most one electron,
  which keeps its distance to some ion inside the cutoff, and the number of such walkers is the
  accepted-move count.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from qmcpack_amd.system import make_system, make_walkers

from test_oracle_nlpp import channels, icosahedron


def test_select_move_known_answers(kats):
    k = kats["tmove_select"]
    for case in k["cases"]:
        prob = math.nextafter(1.0, 0.0) if isinstance(case["prob"], str) else case["prob"]
        i = orc.select_move(prob, k["weight"], k["tau"])
        got = None if i is None else k["pid"][i]
        assert got == case["pid"], case


def test_select_move_against_cumulative_sum():
    rng = np.random.default_rng(3)
    for trial in range(300):
        n = int(rng.integers(0, 40))
        w = rng.normal(size=n) * rng.uniform(0.1, 5)
        w[rng.uniform(size=n) < 0.2] = 0.0
        tau = float(rng.uniform(0.005, 0.5))
        prob = float(rng.uniform())
        f = np.where(w > 0, 0.0, -tau * w)
        scan = 1.0 + np.concatenate([[0.0], np.cumsum(f)[:-1]]) if n else np.zeros(0)
        target = prob * (1.0 + f.sum())
        idx = int(np.searchsorted(scan, target, side="right")) - 1
        expect = None if idx < 0 else idx
        assert orc.select_move(prob, w, tau) == expect, trial


def min_image(d, L):
    return d - L * np.round(d / L)


@pytest.mark.parametrize("seed", [61, 62])
def test_tmoves_follow_the_local_energy(seed):
    s = make_system(L=7.0, n_ion=4, n_up=7, n_dn=6, norb=8, M=12, delay_rank=1, seed=29)
    R = make_walkers(s, 32, 7)
    zion = s.N / len(s.ions)
    ls = [0, 1, 2]
    dr, v, v2, _ = channels(ls, seed=4)
    knots, w = icosahedron()
    rcut = 2.2
    out = []
    for tm in (0, 1):
        o = orc.VMC(s, R, 1, seed)
        o.dmc_set_nlpp(ls, dr, v, v2, rcut, knots, w)
        o.dmc_init(zion, tau=0.3, capacity=128, seed=seed, warmup_steps=3, nonlocal_moves=tm)
        st = o.dmc_run(1)
        out.append((st, o.get()["R"], o.dmc_walker_data()))
    (st0, R0, (w0, e0)), (st1, R1, (w1, e1)) = out
    assert np.array_equal(st0[:, :10], st1[:, :10]) and st0[0, 10] == 0
    assert np.array_equal(w0, w1) and np.array_equal(e0, e1)
    moved = 0
    for iw in range(R.shape[0]):
        diff = np.nonzero(np.any(R0[iw] != R1[iw], axis=1))[0]
        assert len(diff) <= 1
        if len(diff):
            moved += 1
            e = diff[0]
            d0 = np.linalg.norm(min_image(R0[iw, e] - s.ions, s.L), axis=1)
            d1 = np.linalg.norm(min_image(R1[iw, e] - s.ions, s.L), axis=1)
            on_sphere = (np.abs(d0 - d1) < 1e-10) & (d0 < rcut)
            assert np.any(on_sphere), (iw, e, d0, d1)
    print("T-moves accepted", st1[0, 10], "walkers moved", moved)
    assert moved == st1[0, 10] and moved >= 3
