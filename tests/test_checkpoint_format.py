"""The restart-file container (qmi_checkpoint_save / qmi_checkpoint_load, include/qmcmi.h) as read
and written by qmcpack_amd.qmi: the byte layout of the header and of one dataset, a round trip of
every dataset type, and the refusals of malformed input."""
import struct

import numpy as np
import pytest

from qmcpack_amd import qmi


def test_byte_layout():
    data = qmi.write_checkpoint({"state_0/block": np.array([7], dtype=np.int64)})
    name = b"state_0/block"
    expect = (b"QMICKPT1" + struct.pack("<II", 1, 1) + struct.pack("<I", len(name)) + name +
              struct.pack("<II", 2, 1) + struct.pack("<Q", 1) + struct.pack("<q", 7))
    assert data == expect


def test_round_trip():
    rng = np.random.default_rng(0)
    sets = {"state_0/walkers": rng.normal(size=(3, 5, 3)),
            "state_0/number_of_walkers": np.array([3], dtype=np.int64),
            "random/std::mt19937": rng.integers(0, 2 ** 32, size=(2, 625), dtype=np.uint32),
            "branch/history": np.zeros((0, 3))}
    back = qmi.read_checkpoint(qmi.write_checkpoint(sets))
    assert list(back) == list(sets)
    for k, v in sets.items():
        assert back[k].dtype == v.dtype and back[k].shape == v.shape and np.array_equal(back[k], v)


@pytest.mark.parametrize("cut,msg", [(4, "not a walker checkpoint"), (-3, "truncated")])
def test_malformed(cut, msg):
    data = qmi.write_checkpoint({"x": np.arange(4.0)})
    with pytest.raises(ValueError, match=msg):
        qmi.read_checkpoint(data[:cut])
    with pytest.raises(ValueError, match="trailing"):
        qmi.read_checkpoint(data + b"\0")


def test_oracle_stream_loader_against_numpy():
    """The oracle's StdRandom::load restatement (used to start the oracle from a checkpoint's
    streams): numpy's MT19937 states of fresh seeds 77 + c, loaded into an oracle seeded with 5,
    give the run of an oracle seeded with 77, bit for bit."""
    from oracle import oracle as orc
    from qmcpack_amd.system import make_system, make_walkers
    s = make_system(L=7.0, n_ion=3, n_up=7, n_dn=6, norb=8, M=12, delay_rank=4, seed=13)
    R = make_walkers(s, 4, 5)
    a = orc.VMC(s, R, 2, 77)
    a.evaluate_log()
    Ea = a.run(2, 1, 0.4, 0)[0]
    b = orc.VMC(s, R, 2, 5)
    states = np.zeros((2, 625), dtype=np.uint32)
    for c in range(2):
        st = np.random.RandomState(77 + c)._bit_generator.state["state"]
        states[c, :624], states[c, 624] = st["key"], st["pos"]
    b.set_rng_states(states)
    b.evaluate_log()
    Eb = b.run(2, 1, 0.4, 0)[0]
    assert np.array_equal(Ea, Eb) and np.array_equal(a.get()["R"], b.get()["R"])
    with pytest.raises(RuntimeError):
        b.set_rng_states(states[:1])
