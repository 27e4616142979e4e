"""The oracle's restatement of the dense SoA distance tables (SoaDistanceTableAA / AB), pinned by the
reference's known-answer sequence of test_distance_table.cpp (proposal rows, old rows, the table
after a regular accept, forward mode with updatePartial) and by what the mathematics fixes: after any
sequence of regular-mode moves the maintained part equals the from-scratch table of the current
positions, and so does the lower triangle after a forward sweep over all electrons.
"""
import numpy as np
import pytest

from oracle import oracle as orc

REL = 2e-9  # reference values carry ten digits


def close(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.all(np.abs(a - b) <= REL * np.abs(b) + 1e-12)


def test_known_answer_sequence(kats):
    k = kats["distance_table_sequence"]
    L = [k["L"]] * 3
    R = np.array(k["elec"], dtype=float)
    t = orc.DistanceTable(L, "AA", 3)
    t.evaluate(R)
    shift, disp = np.array(k["shift_e0"]), np.array(k["disp"])
    t.move(R, R[0] + shift, 0)
    t.update(0)
    R[0] += shift
    assert close(t.r[1, 0], k["after_shift_row1_col0"][0]) and close(t.dr[1, 0], k["after_shift_row1_col0"][1])
    t.move(R, R[0] + disp, 0, prepare_old=False)  # then rejected: no update
    assert close(t.temp_r[1], k["move0_temp1"][0]) and close(t.temp_dr[1], k["move0_temp1"][1])
    assert close(t.r[1, 0], k["after_shift_row1_col0"][0])
    t.move(R, R[0] + disp, 0)
    assert close(t.old_r[1], k["move0_old1"][0]) and close(t.old_dr[1], k["move0_old1"][1])
    assert t.old_r[0] == np.finfo(float).max
    t.update(0)
    R[0] += disp
    assert close(t.r[1, 0], k["after_accept_row1_col0"][0]) and close(t.dr[1, 0], k["after_accept_row1_col0"][1])
    t.move(R, R[0] - disp, 0)
    t.update(0)
    R[0] -= disp
    # forward mode: accepting electron 0 writes row 0 only, so row 1 keeps its old entry
    t.move(R, R[0] + disp, 0)
    t.update_partial(0, True)
    R[0] += disp
    assert close(t.r[1, 0], k["forward_after_e0_row1_col0"][0])
    assert close(t.dr[1, 0], k["forward_after_e0_row1_col0"][1])
    # electron 1's rejected move writes row 1 from its old rows, which see the moved electron 0
    t.move(R, R[1] + disp, 1)
    t.update_partial(1, False)
    assert close(t.r[1, 0], k["forward_after_e1_rejected_row1_col0"][0])
    assert close(t.dr[1, 0], k["forward_after_e1_rejected_row1_col0"][1])


def _scratch(L, kind, R, ions=None):
    t = orc.DistanceTable(L, kind, R.shape[0], ions)
    t.evaluate(R)
    return t


@pytest.mark.parametrize("kind", ["AA", "AB"])
def test_regular_mode_equals_from_scratch(kind):
    rng = np.random.default_rng(1)
    L = np.array([5.0, 6.0, 7.0])
    N = 9
    R = rng.uniform(0, 1, size=(N, 3)) * L
    ions = rng.uniform(0, 1, size=(4, 3)) * L
    t = _scratch(L, kind, R, ions)
    for _ in range(60):
        iat = int(rng.integers(N))
        rnew = R[iat] + rng.normal(size=3)
        t.move(R, rnew, iat)
        if rng.integers(2):
            t.update(iat)
            R[iat] = rnew
        ref = _scratch(L, kind, R, ions)
        m = ~np.isnan(ref.r)
        assert np.array_equal(m, ~np.isnan(t.r))
        assert np.allclose(t.r[m], ref.r[m], rtol=1e-14, atol=1e-14)
        assert np.allclose(t.dr[m], ref.dr[m], rtol=1e-14, atol=1e-14)


def test_forward_sweep_equals_from_scratch():
    rng = np.random.default_rng(2)
    L = np.array([6.0, 6.0, 6.0])
    N = 8
    R = rng.uniform(0, 1, size=(N, 3)) * L
    t = _scratch(L, "AA", R)
    for iat in range(N):
        rnew = R[iat] + rng.normal(size=3)
        t.move(R, rnew, iat)
        acc = bool(rng.integers(2))
        t.update_partial(iat, acc)
        if acc:
            R[iat] = rnew
    ref = _scratch(L, "AA", R)
    m = ~np.isnan(ref.r)
    assert np.allclose(t.r[m], ref.r[m], rtol=1e-14, atol=1e-14)
    assert np.allclose(t.dr[m], ref.dr[m], rtol=1e-14, atol=1e-14)


def test_skipping_the_column_update_is_caught():
    """A regular update that writes only the row (as forward mode does) leaves stale entries below
    the diagonal: the from-scratch comparison above sees it."""
    rng = np.random.default_rng(3)
    L = np.array([6.0, 6.0, 6.0])
    R = rng.uniform(0, 1, size=(5, 3)) * L
    t = _scratch(L, "AA", R)
    rnew = R[1] + np.array([0.7, -0.3, 0.2])
    t.move(R, rnew, 1)
    t.update_partial(1, True)  # row only
    R[1] = rnew
    ref = _scratch(L, "AA", R)
    m = ~np.isnan(ref.r)
    assert not np.allclose(t.r[m], ref.r[m], rtol=1e-14, atol=1e-14)
