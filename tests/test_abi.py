"""The C-ABI library loads and exports every entry point declared in include/qmcmi.h.
No compute calls: this runs without a GPU."""
import ctypes as C

from qmcpack_amd import qmi


def test_header_declares_entry_points():
    names = qmi.exported_symbols_from_header()
    for must in ("qmi_ctx_create", "qmi_mw_eval_grad", "qmi_mw_ratio_grad", "qmi_mw_accept_reject",
                 "qmi_mw_complete_updates", "qmi_mw_evaluate_gl", "qmi_mw_evaluate_log", "qmi_vmc_run"):
        assert must in names


def test_library_exports_all_symbols():
    lib = qmi.load_library()
    missing = [n for n in qmi.exported_symbols_from_header() if not hasattr(lib, n)]
    assert not missing, missing


def test_system_desc_layout_matches_oracle():
    from oracle.oracle import OrcSystemDesc
    assert C.sizeof(qmi.SystemDesc) == C.sizeof(OrcSystemDesc)
    for (a, ta), (b, tb) in zip(qmi.SystemDesc._fields_, OrcSystemDesc._fields_):
        assert a == b and C.sizeof(ta) == C.sizeof(tb)
        assert getattr(qmi.SystemDesc, a).offset == getattr(OrcSystemDesc, b).offset
