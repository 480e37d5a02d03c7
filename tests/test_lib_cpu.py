"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/trans_bl_amd.h declares; host-only routines run."""
import ctypes
import os
import re

import numpy as np
import pytest

import trans_bl_amd as tb

HDR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "trans_bl_amd.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tbl_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_python_export_list():
    assert declared() == sorted(tb.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tb.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_text():
    assert b"gfx950" in tb.lib().tbl_version()
    assert isinstance(tb.lib().tbl_last_error(), bytes)


def test_coefs_generate_rejects_bad_arguments():
    c = tb.CCoefs()
    assert tb.lib().tbl_coefs_generate(1.01, 0, ctypes.byref(c)) == 1
    assert tb.lib().tbl_coefs_generate(-1.0, 4, ctypes.byref(c)) == 1
    assert tb.lib().tbl_coefs_generate(1.01, 4, None) == 1


def test_null_context_is_rejected():
    L = tb.lib()
    assert L.tbl_destroy(None) == 1
    assert L.tbl_synchronize(None) == 1
    assert L.tbl_deriv_dev(None, 0, None, None) == 1


def test_coefs_roundtrip_through_c_struct():
    d = tb.coefs_generate(1.02, 4)
    back = tb.coefs_to_dict(tb.coefs_from_dict(d), 4)
    for k in d:
        np.testing.assert_array_equal(d[k], back[k])


def test_coefs_bin_reader(tmp_path):
    """The reader takes the reference's record layout; the shipped fixture
    has the stale 7-entry first record and is rejected, a file with the
    current 6-entry record is read back exactly."""
    import struct
    from oracle import coefgen
    fix = os.path.join(os.path.dirname(__file__), "golden", "coefs_stf1.01_lvl4.bin")
    with pytest.raises(tb.TblError):
        tb.coefs_read(fix, 4)
    recs = coefgen.read_coefs_bin(fix)
    g = coefgen.generate(1.01, 4)
    recs[0] = g["fp_fd_e"]
    path = tmp_path / "coefs.bin"
    with open(path, "wb") as f:
        for r in recs:
            b = np.asarray(r, dtype="<f8").tobytes()
            f.write(struct.pack("<i", len(b)) + b + struct.pack("<i", len(b)))
    d = tb.coefs_read(str(path), 4)
    np.testing.assert_array_equal(d["fp_fd_e"], g["fp_fd_e"])
    for i, k in enumerate(["sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd", "pp_sd",
                           "lp_sd"]):
        np.testing.assert_array_equal(d[k], recs[i + 1])
    for i, k in enumerate(["sp_poi", "cp_poi", "pp_poi", "lp_poi"]):
        np.testing.assert_array_equal(d[k].ravel(), recs[10 + i])
    with pytest.raises(tb.TblError):
        tb.coefs_read(str(path), 3)        # wrong level count
