"""The allocation-free inverter (qmi_mw_invert_transpose_noalloc) against qmi_mw_invert_transpose.

Both run the same kernels with the same arithmetic, so the inverses and log determinants must agree
bit for bit.  Two calls in a row on one context, the second with fewer and smaller matrices (the
scratch keeps the first call's contents, which must not leak into the second), for real and complex
matrices; a singular matrix is reported through info where the allocating entry point raises; a call
beyond the scratch is refused before any work.  Device buffers come from torch (device memory only;
the library's stream is separate, so the test synchronises the device around each call).
"""
import numpy as np
import pytest
import torch

from qmcpack_amd import qmi

pytestmark = pytest.mark.gpu

QMI_ERR_ARG, QMI_ERR_NUMERIC = 1, 3


def matrices(rng, nmat, n, cplx):
    a = rng.standard_normal((nmat, n, n)) + n ** 0.5 * np.eye(n)
    if cplx:
        a = a + 1j * rng.standard_normal((nmat, n, n))
    return a


def as_pairs(a):
    return np.ascontiguousarray(np.stack([a.real, a.imag], -1) if np.iscomplexobj(a) else a, dtype=np.float64)


def run_noalloc(c, a):
    cplx = np.iscomplexobj(a)
    nmat, n = a.shape[0], a.shape[1]
    da = torch.from_numpy(as_pairs(a)).cuda()
    dout = torch.full_like(da, float("nan"))
    dlog = torch.full((nmat, 2), float("nan"), dtype=torch.float64, device="cuda")
    dinfo = torch.full((nmat,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    c.mw_invertTranspose_noalloc(cplx, n, nmat, da.data_ptr(), dout.data_ptr(), dlog.data_ptr(), dinfo.data_ptr())
    torch.cuda.synchronize()
    out, lv, info = dout.cpu().numpy(), dlog.cpu().numpy(), dinfo.cpu().numpy()
    inv = out[..., 0] + 1j * out[..., 1] if cplx else out
    return inv, lv[:, 0] + 1j * lv[:, 1], info


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_matches_allocating_entry(cplx):
    rng = np.random.default_rng(71 + cplx)
    c = qmi.Context(0)
    for nmat, n in ((6, 45), (3, 17)):  # the second call is smaller
        a = matrices(rng, nmat, n, cplx)
        inv_ref, lv_ref = c.mw_invertTranspose(a)
        inv, lv, info = run_noalloc(c, a)
        print(f"nmat {nmat} n {n} complex {cplx}: max |diff| {np.max(np.abs(inv - inv_ref)):.1e}")
        assert np.array_equal(info, np.zeros(nmat, dtype=np.int32))
        assert np.array_equal(inv, inv_ref)
        assert np.array_equal(lv, lv_ref)
        # and it is the inverse transposed
        assert np.allclose(np.einsum("mij,mkj->mik", a, inv), np.eye(n), atol=1e-10)


def test_noalloc_singular_and_limits():
    rng = np.random.default_rng(73)
    c = qmi.Context(0)
    a = matrices(rng, 3, 12, False)
    a[1, :, 4] = 0.0  # column 4 of matrix 1 is zero: A(i, 4) = 0 for every i
    with pytest.raises(qmi.QmiError) as e:
        c.mw_invertTranspose(a)
    assert e.value.code == QMI_ERR_NUMERIC
    inv, lv, info = run_noalloc(c, a)
    assert info[0] == 0 and info[2] == 0 and info[1] == 5
    inv_ok, lv_ok = c.mw_invertTranspose(a[[0, 2]])
    assert np.array_equal(inv[[0, 2]], inv_ok) and np.array_equal(lv[[0, 2]], lv_ok)
    # more than the scratch holds: 64 matrices of 512 x 512 need 2 x 128 MiB
    with pytest.raises(qmi.QmiError) as e:
        c.mw_invertTranspose_noalloc(False, 512, 64, 1, 1, 0, 1)
    assert e.value.code == QMI_ERR_ARG and "scratch" in str(e.value)
