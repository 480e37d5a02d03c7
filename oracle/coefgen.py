"""Oracle restatement of the reference's y-direction coefficient generator.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import it, and only as the reference that results are
compared against.

What it restates
----------------
``pre_processing/derivative.f90`` builds, for every compact scheme on the
geometric grid, a small linear system of Taylor-moment conditions and solves it
with a no-pivot Doolittle LU (``pre_processing/solvers.f90:73-194``), zeroing
any entry whose magnitude is below 1e-14 (solvers.f90:126, 189).  The
derivative (fd/sd) schemes are solved in real*8 at the grid stretch ``stf``
(derivative.f90:1762-1845); the Poisson schemes in real(kind=10) at the level
stretch ``stf**(2**(l-1))`` (derivative.f90:1850-1882,
coef_generation.f90:32-38).

This module does not copy those matrices.  Each scheme is described here by
its unknowns -- (side, derivative order, normalised offset) -- and by its
fixed entries; the moment conditions are generated from that description:

    for p = 0..P:  sum_u side_u * c_u * h_u**(p - q_u) / (p - q_u)!  = 0
                   (terms with p < q_u vanish)

solved in 80-bit ``numpy.longdouble`` with partial pivoting, flushed at 1e-14
and rounded to float64.  A non-singular system has one solution, so the row
order and sign conventions of the reference's matrices do not matter; the
values differ from the reference's real*8 solves only by their rounding
(about 1e-15 relative, see the coefs.bin test).

The coefficient index order of every output array is the interface: it is the
order of ``COMMON /coefs/`` (comm.coef:1-29) and of ``coefs.bin``
(coef_generation.f90:39-60), which the product and the reference's readers
index directly.

Normalised offsets: the grid interval that the reference's row scaling uses
(``v_qdy(j-2)``, ``v_qdy(1)``, ``v_qdy(jmax-5)`` ...) has length 1 and each
following interval is ``s`` times the previous one (derivs_par.f:400-611,
poisson_par.f:229-283).
"""
from __future__ import annotations

import math
import numpy as np

LD = np.longdouble
FLUSH = 1e-14

# unknown kinds
F, D1, D2 = 0, 1, 2
LHS, RHS = +1, -1


def _solve_ld(m, b):
    """Gaussian elimination with partial pivoting in 80-bit long double."""
    a = np.array(m, dtype=LD)
    x = np.array(b, dtype=LD)
    n = len(x)
    for c in range(n):
        p = c + int(np.argmax(np.abs(a[c:, c])))
        if a[p, c] == 0:
            raise ValueError("singular moment system")
        if p != c:
            a[[c, p]] = a[[p, c]]
            x[[c, p]] = x[[p, c]]
        for r in range(c + 1, n):
            f = a[r, c] / a[c, c]
            if f != 0:
                a[r, c:] -= f * a[c, c:]
                x[r] -= f * x[c]
    out = np.zeros(n, dtype=LD)
    for r in range(n - 1, -1, -1):
        out[r] = (x[r] - np.dot(a[r, r + 1:], out[r + 1:])) / a[r, r]
    return out


def solve_scheme(unknowns, fixed, pmax):
    """unknowns: list of (side, q, h); fixed: {index: value}; moments 0..pmax."""
    n = len(unknowns)
    rows, rhs = [], []
    for p in range(pmax + 1):
        row = []
        for side, q, h in unknowns:
            if p < q:
                row.append(LD(0))
            else:
                e = p - q
                row.append(LD(side) * LD(h) ** e / LD(math.factorial(e)))
        rows.append(row)
        rhs.append(LD(0))
    for idx, val in fixed.items():
        row = [LD(0)] * n
        row[idx] = LD(1)
        rows.append(row)
        rhs.append(LD(val))
    if len(rows) != n:
        raise ValueError(f"scheme has {n} unknowns but {len(rows)} conditions")
    x = _solve_ld(rows, rhs)
    x[np.abs(x) < FLUSH] = 0
    return x


def _cum(steps):
    """Positions reached by walking the listed interval lengths (long double)."""
    out, acc = [], LD(0)
    for s in steps:
        acc = acc + s
        out.append(acc)
    return out


# ---------------------------------------------------------------------------
# first-derivative (fd) schemes: derivative.f90:15-612, 2002-2081
# ---------------------------------------------------------------------------
def fd_schemes(stf):
    s = LD(stf)
    # point 1 at the wall; upward intervals 1, s, s^2, s^3
    up = [LD(0)] + _cum([LD(1), s, s**2, s**3])
    fp_e = solve_scheme([(LHS, D1, 0)] + [(RHS, F, h) for h in up], {0: 12}, 4)
    fp = solve_scheme([(LHS, D1, 0), (LHS, D1, up[1])] + [(RHS, F, h) for h in up],
                      {0: 1}, 5)
    # point 2: one interval (length 1) below, then s, s^2, s^3, s^4 above
    abv = _cum([s, s**2, s**3, s**4])
    pts = [LD(-1), LD(0)] + abv
    sp = solve_scheme([(LHS, D1, -1), (LHS, D1, 0), (LHS, D1, abv[0])]
                      + [(RHS, F, h) for h in pts], {0: 1, 4: -2.5}, 6)
    # centre point j: intervals (j-2,j-1)=1, (j-1,j)=s, (j,j+1)=s^2, (j+1,j+2)=s^3
    cpts = [-(LD(1) + s), -s, LD(0), s**2, s**2 + s**3]
    cp = solve_scheme([(LHS, D1, cpts[1]), (LHS, D1, 0), (LHS, D1, cpts[3])]
                      + [(RHS, F, h) for h in cpts], {1: 1}, 6)
    # penultimate point jmax-1: (jmax-5,jmax-4)=1 ... (jmax-1,jmax)=s^4
    below = [-h for h in _cum([s**3, s**2, s, LD(1)])]
    ppts = [s**4, LD(0)] + below          # f(jmax), f(jmax-1), f(jmax-2..jmax-5)
    pp = solve_scheme([(LHS, D1, s**4), (LHS, D1, 0), (LHS, D1, below[0])]
                      + [(RHS, F, h) for h in ppts], {0: 1, 4: 2.5}, 6)
    # last point jmax: (jmax-4,jmax-3)=1 ... (jmax-1,jmax)=s^3
    lbelow = [-h for h in _cum([s**3, s**2, s, LD(1)])]
    lpts = [LD(0)] + lbelow
    lp = solve_scheme([(LHS, D1, 0), (LHS, D1, lbelow[0])]
                      + [(RHS, F, h) for h in lpts], {0: 1}, 5)
    return dict(fp_fd=fp, fp_fd_e=fp_e, sp_fd=sp, cp_fd=cp, pp_fd=pp, lp_fd=lp)


# ---------------------------------------------------------------------------
# second-derivative (sd) schemes: derivative.f90:617-1267
# ---------------------------------------------------------------------------
def sd_schemes(stf):
    s = LD(stf)
    up = [LD(0)] + _cum([LD(1), s, s**2, s**3, s**4])
    fp = solve_scheme([(LHS, D2, 0), (LHS, D2, up[1])] + [(RHS, F, h) for h in up],
                      {0: 1}, 6)
    abv = _cum([s, s**2, s**3, s**4, s**5])
    pts = [LD(-1), LD(0)] + abv
    sp = solve_scheme([(LHS, D2, -1), (LHS, D2, 0), (LHS, D2, abv[0])]
                      + [(RHS, F, h) for h in pts], {0: 1, 5: 5.25}, 7)
    cpts = [-(LD(1) + s), -s, LD(0), s**2, s**2 + s**3]
    cp = solve_scheme([(LHS, D2, cpts[1]), (LHS, D2, 0), (LHS, D2, cpts[3])]
                      + [(RHS, F, h) for h in cpts], {1: 1}, 6)
    below = [-h for h in _cum([s**4, s**3, s**2, s, LD(1)])]
    ppts = [s**5, LD(0)] + below
    pp = solve_scheme([(LHS, D2, s**5), (LHS, D2, 0), (LHS, D2, below[0])]
                      + [(RHS, F, h) for h in ppts], {0: 1, 5: 5.25}, 7)
    lbelow = [-h for h in _cum([s**4, s**3, s**2, s, LD(1)])]
    lpts = [LD(0)] + lbelow
    lp = solve_scheme([(LHS, D2, 0), (LHS, D2, lbelow[0])]
                      + [(RHS, F, h) for h in lpts], {0: 1}, 6)
    return dict(fp_sd=fp, sp_sd=sp, cp_sd=cp, pp_sd=pp, lp_sd=lp)


# ---------------------------------------------------------------------------
# Poisson schemes, one set per multigrid level: derivative.f90:1272-1613
# ---------------------------------------------------------------------------
def poisson_schemes(strl):
    s = LD(strl)
    # row 2: f'_1 (wall) + B-weights on f''_2, f''_3 ; f at points 1..4
    pts = [LD(-1), LD(0), s, s + s**2]
    sp = solve_scheme([(LHS, D1, -1), (LHS, D2, 0), (LHS, D2, s)]
                      + [(RHS, F, h) for h in pts], {1: LD(4) / LD(5)}, 5)
    cpts = [-(LD(1) + s), -s, LD(0), s**2, s**2 + s**3]
    cp = solve_scheme([(LHS, D2, -s), (LHS, D2, 0), (LHS, D2, s**2)]
                      + [(RHS, F, h) for h in cpts], {1: LD(11) / LD(15)}, 6)
    # row ny-1: (ny-2,ny-1)=1, (ny-1,ny)=s
    pp = solve_scheme([(LHS, D2, 0), (RHS, F, -1), (RHS, F, 0), (RHS, F, s)],
                      {0: 1}, 2)
    # row ny: f''_ny = lp3 f_ny + lp4 f_ny-1 + lp5 f_ny-2 + lp2 f'_ny
    lp = solve_scheme([(LHS, D2, 0), (RHS, D1, 0), (RHS, F, 0), (RHS, F, -s),
                       (RHS, F, -(s + LD(1)))], {0: 1}, 3)
    return dict(sp_poi=sp, cp_poi=cp, pp_poi=pp, lp_poi=lp)


# ---------------------------------------------------------------------------
# wall tables: derivative.f90:1618-1757 (w_at_w), 1887-1998 (dwydy), solved in
# real(kind=10) (derivative.f90:1838-1843)
# ---------------------------------------------------------------------------
def wall_schemes(stf):
    s = LD(stf)
    hp = _cum([LD(1), s, s**2, s**3, s**4, s**5])     # rows 2..7 above the wall
    # dwydy: f'(wall) = sum_j c(j+1) f(j), j = 1..7, with c(1) = 1
    dwydy = solve_scheme([(LHS, D1, 0)] + [(RHS, F, h) for h in [LD(0)] + hp], {0: 1}, 6)
    # w_at_w: c(1) f''(wall) + c(2) f'(wall) = -sum_j c(j+2) f(j), exact to the
    # 7th moment, normalised by c(9) = -1/18.  The reference's first-moment row
    # also carries -c(3) (derivative.f90:1645-1653), which moves only c(2);
    # c(2) enters no formula of the path.
    w = solve_scheme([(LHS, D2, 0), (LHS, D1, 0)] + [(RHS, F, h) for h in [LD(0)] + hp],
                     {8: LD(-1) / LD(18)}, 7)
    w[1] = w[2] + sum(w[3 + n] * hp[n] for n in range(6))
    if abs(w[1]) < FLUSH:
        w[1] = 0
    return dict(w_at_w=w, dwydy=dwydy)


def level_stretch(stf, lvl):
    """stf**(2**(lvl-1)) as the reference evaluates it (real*8 power)."""
    return float(stf) ** (2 ** (lvl - 1))


def generate(stf, levels):
    """All tables for a grid stretch and a level count, rounded to float64.

    Returns a dict whose keys follow COMMON /coefs/; the Poisson tables are
    arrays of shape (levels, ncoef).
    """
    out = {}
    for k, v in fd_schemes(stf).items():
        out[k] = v.astype(np.float64)
    for k, v in sd_schemes(stf).items():
        out[k] = v.astype(np.float64)
    for k, v in wall_schemes(stf).items():
        out[k] = v.astype(np.float64)
    per = [poisson_schemes(level_stretch(stf, l)) for l in range(1, levels + 1)]
    for k in ("sp_poi", "cp_poi", "pp_poi", "lp_poi"):
        out[k] = np.stack([p[k].astype(np.float64) for p in per])
    return out


def read_coefs_bin(path):
    """Decode a coefs.bin (20 sequential unformatted records, 4-byte markers;
    coef_generation.f90:39-60).  Returns the list of float64 records."""
    import struct
    data = open(path, "rb").read()
    off, recs = 0, []
    while off < len(data):
        n = struct.unpack("<i", data[off:off + 4])[0]
        recs.append(np.frombuffer(data[off + 4:off + 4 + n], dtype="<f8").copy())
        tail = struct.unpack("<i", data[off + 4 + n:off + 8 + n])[0]
        if tail != n:
            raise ValueError("corrupt record marker")
        off += 8 + n
    return recs
