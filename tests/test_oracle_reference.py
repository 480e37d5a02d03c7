"""The oracle against the reference's own Fortran (golden vectors).

tests/golden/ref_vectors.npz holds outputs of the reference's poi_uy, its
five derivatives and its coefficient generator, compiled from the reference
sources and run by oracle/reference.py (script: tests/golden/
make_ref_vectors.py).  The inputs are the seeded wavepackets of
trans_bl_amd.workload, regenerated here.  This pins what no closed form
does: the exact iterate sequence of the reference's lexicographic line-SOR
inside its FAS cycle (V-cycle counts at the 1e-9 stop, and the iterate after
11 cycles without it), with nonzero wx.  When the reference drivers are built
in oracle/_ref/ the last test also compares whole fields directly.
"""
import os

import numpy as np
import pytest

import oracle
import trans_bl_amd as tb
from oracle import reference as R
from trans_bl_amd import workload as wl
from conftest import rel

VEC = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "ref_vectors.npz"))
SUB = (slice(None), slice(None, None, 4), slice(None, None, 8))
SUB_BIG = {"385x97x8": (slice(None), slice(None, None, 8), slice(None, None, 16)),
           "1025x193x64": ([0, 1, 31, 63], slice(None, None, 8), slice(None, None, 16))}
TABLES = ("fp_fd_e", "sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd", "pp_sd",
          "lp_sd", "sp_poi", "cp_poi", "pp_poi", "lp_poi")


def grid_of(name):
    nx, ny, nk, msh, dx, dy, stf, alpha, beta = R.CONFIGS[name]
    return wl.Grid(nx, ny, nk, msh, dx, dy, stf, alpha, beta)


def ref_coefs(name):
    return {t: VEC[f"{name}/coefs/{t}"] for t in TABLES}


def params(name):
    g = grid_of(name)
    return g, oracle.make_params(g.nx, g.ny, g.nmodes, g.levels, g.dx, g.dy0, g.stf, g.alpha,
                                 g.beta, coefs=ref_coefs(name))


@pytest.mark.parametrize("name", ["129x33x4", "385x97x8", "1025x193x64"])
def test_generators_match_reference_generator(name):
    """The oracle's and the library's coefficient generators at the
    workloads' stretches (1.03, 1.02, 1.015) against the reference's
    generator run on the same par.for."""
    g = grid_of(name)
    ref = ref_coefs(name)
    mine = oracle.coefgen.generate(g.stf, g.levels)
    lib = tb.coefs_generate(g.stf, g.levels)
    for t in TABLES:
        b = np.asarray(ref[t])
        scale = np.max(np.abs(b))
        assert np.max(np.abs(np.asarray(mine[t]) - b)) < 1e-13 * scale, t
        assert np.max(np.abs(np.asarray(lib[t]) - b)) < 1e-13 * scale, t


@pytest.fixture(scope="module")
def small():
    name = "129x33x4"
    g, p = params(name)
    f = wl.wavepacket_rhs(g, tag="refvec")
    wx = 0.1 * wl.wavepacket_rhs(g, tag="refvec-wx")
    return name, g, p, f, wx


def test_lexicographic_iterate_matches_reference(small):
    """11 V-cycles with no stop (erromax = 0): the iterate of the reference's
    own smoother order, rf schedule, FAS cycle and intpol, bit-close."""
    name, g, p, f, wx = small
    z = np.zeros_like(f)
    v, cyc, _ = oracle.poi_uy(p, wx, -f, z, 0.0, max_cycles=11, order=0)
    assert np.array_equal(cyc, VEC[f"{name}/iter11/cycles"])
    assert rel(v[SUB], VEC[f"{name}/iter11/uy"]) < 1e-13


def test_stop_rule_matches_reference(small):
    name, g, p, f, wx = small
    z = np.zeros_like(f)
    v, cyc, res = oracle.poi_uy(p, wx, -f, z, 1e-9, max_cycles=11, order=0)
    assert np.array_equal(cyc, VEC[f"{name}/stop/cycles"])
    assert rel(v[SUB], VEC[f"{name}/stop/uy"]) < 1e-13
    np.testing.assert_allclose(res, VEC[f"{name}/stop/residual"], rtol=1e-3)


def test_derivatives_match_reference(small):
    name, g, p, _, _ = small
    d = wl.wavepacket_rhs(g, tag="refvec-deriv")
    for kind in ("x", "xx", "y", "yy", "yfv"):
        got = oracle.deriv(p, kind, d)[SUB]
        assert rel(got, VEC[f"{name}/deriv/{kind}"]) < 1e-13, kind


@pytest.mark.parametrize("name, tag, modes", [
    ("385x97x8", "refvec-cfg1", list(range(8))),
    ("1025x193x64", "refvec-cfg2", [0, 1, 31, 63]),
])
def test_benchmark_workloads_match_reference(name, tag, modes):
    """The two benchmark grids from zero to the 1e-9 stop: the reference's
    V-cycle count of every checked mode, and v and its derivatives."""
    g, p = params(name)
    f = wl.wavepacket_rhs(g, tag=tag)
    z = np.zeros_like(f)
    v, cyc, _ = oracle.poi_uy_modes(p, z, -f, z, 1e-9, modes, max_cycles=11, order=0)
    assert np.array_equal(cyc[modes], VEC[f"{name}/stop/cycles"][modes])
    sub = SUB_BIG[name]
    sel = (modes,) + sub[1:]
    assert rel(v[sel], VEC[f"{name}/stop/uy"]) < 1e-12
    for kind in ("x", "xx", "y", "yy"):
        d = oracle.deriv(p, kind, v)[sel]
        assert rel(d, VEC[f"{name}/stop/{kind}"]) < 1e-12, kind


@pytest.mark.skipif(not R.available("129x33x4"), reason="reference driver not built")
def test_whole_fields_against_reference_driver(small):
    """Whole fields: the oracle and the reference driver in this process."""
    name, g, p, f, wx = small
    z = np.zeros_like(f)
    r = R.run(name, wx, -f, z, 0.0)
    v, _, _ = oracle.poi_uy(p, wx, -f, z, 0.0, max_cycles=11, order=0)
    assert rel(v, r["uy"]) < 1e-13
    for kind in ("x", "xx", "y", "yy", "yfv"):
        assert rel(oracle.deriv(p, kind, r["uy"]), r[kind]) < 1e-13, kind
