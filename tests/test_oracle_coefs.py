"""Pins of the coefficient generators against the reference's only known
answer, pre_processing/coefs.bin (stf = 1.01, 4 levels), and the closed forms
at stf = 1 (Lele's alpha = 1/3 and 2/11 compact schemes; the Poisson schemes
of derivative.f90:1272-1613 at unit stretch)."""
import math
import os

import numpy as np
import pytest

from oracle import coefgen

FIX = os.path.join(os.path.dirname(__file__), "golden", "coefs_stf1.01_lvl4.bin")

# record order of coefs.bin (coef_generation.f90:39-60); record 1 in the
# shipped file is the older implicit first-point scheme (7 entries)
RECORDS = ["fp_fd", "sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd", "pp_sd",
           "lp_sd", "sp_poi", "cp_poi", "pp_poi", "lp_poi"]


@pytest.fixture(scope="module")
def recs():
    return coefgen.read_coefs_bin(FIX)


def test_fixture_layout(recs):
    assert [len(r) for r in recs] == [7, 9, 8, 9, 7, 8, 10, 8, 10, 8, 28, 32, 16, 20, 9, 8, 7, 8, 7, 7]
    # the stretch the fixture was made for: pp_poi centre = -2/strl at each level
    pp = recs[12].reshape(4, 4)
    strl = -2.0 / pp[:, 2]
    assert np.allclose(strl, [1.01, 1.01 ** 2, 1.01 ** 4, 1.01 ** 8], rtol=1e-14)


@pytest.mark.parametrize("i,name", list(enumerate(RECORDS)))
def test_oracle_generator_matches_coefs_bin(recs, i, name):
    g = coefgen.generate(1.01, 4)
    ref = recs[i]
    got = np.asarray(g[name]).ravel()
    assert got.shape == ref.shape
    # Poisson tables: both sides solve in 80-bit; derivative tables: the
    # reference solved in real*8 (its own rounding is up to ~1e-13)
    tol = 2e-15 if name.endswith("_poi") else 2e-13
    np.testing.assert_allclose(got, ref, rtol=tol, atol=1e-15)


def test_closed_forms_unit_stretch():
    g = coefgen.generate(1.0, 2)
    np.testing.assert_allclose(g["cp_fd"], [1 / 3, 1, 1 / 3, -1 / 36, -7 / 9, 0, 7 / 9, 1 / 36],
                               rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(g["cp_sd"], [2 / 11, 1, 2 / 11, 3 / 44, 12 / 11, -51 / 22, 12 / 11,
                                            3 / 44], rtol=1e-15)
    for l in range(2):
        s = 1.0 ** (2 ** l)
        np.testing.assert_allclose(g["cp_poi"][l], [2 / 15, 11 / 15, 2 / 15, 1 / 20, 4 / 5,
                                                    -17 / 10, 4 / 5, 1 / 20], rtol=1e-15)
        np.testing.assert_allclose(g["pp_poi"][l], [1, 1, -2, 1], rtol=1e-15)
        np.testing.assert_allclose(g["lp_poi"][l], [1, 3, -3.5, 4, -0.5], rtol=1e-15)
        assert g["sp_poi"][l][1] == 0.8
    # the first-point explicit scheme is the classical one-sided 4th-order
    # formula 12 f' = -25 f0 + 48 f1 - 36 f2 + 16 f3 - 3 f4
    np.testing.assert_allclose(g["fp_fd_e"], [12, -25, 48, -36, 16, -3], rtol=1e-14)


def _moment_residual(weights_lhs, weights_rhs, q, pmax):
    out = []
    for p in range(pmax + 1):
        lhs = sum(c * (h ** (p - q) / math.factorial(p - q) if p >= q else 0.0)
                  for c, h in weights_lhs)
        rhs = sum(c * h ** p / math.factorial(p) for c, h in weights_rhs)
        out.append(lhs - rhs)
    return np.array(out)


def test_centre_schemes_are_moment_exact_on_stretched_grid():
    s = 1.02
    g = coefgen.generate(s, 1)
    pts = [-(1 + s), -s, 0.0, s * s, s * s + s ** 3]
    lhs_off = [-s, 0.0, s * s]
    for name, q in (("cp_fd", 1), ("cp_sd", 2), ("cp_poi", 2)):
        c = np.asarray(g[name]).ravel()
        r = _moment_residual(list(zip(c[:3], lhs_off)), list(zip(c[3:8], pts)), q, 6)
        assert np.max(np.abs(r)) < 1e-12, (name, r)


def test_library_generator_matches_oracle_generator():
    import trans_bl_amd as tb
    for stf, lv in ((1.01, 4), (1.015, 4), (1.02, 4), (1.0, 3)):
        a = tb.coefs_generate(stf, lv)
        b = coefgen.generate(stf, lv)
        for k in b:
            if k == "fp_fd":
                continue
            np.testing.assert_allclose(np.asarray(a[k]).ravel(), np.asarray(b[k]).ravel(),
                                       rtol=1e-14, atol=1e-15, err_msg=f"{k} stf={stf}")
