// Host-side setup of the trans_bl_amd library: coefficient generation,
// multigrid metrics, line-LU factors and Thomas factors.  Plain C++; runs once
// per context (the reference does this offline or in init_gv).
#include "tbl_host.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace tbl {

namespace {

using ld = long double;

enum Kind { KF = 0, KD1 = 1, KD2 = 2 };

struct Unknown {
    int side;  // +1 left-hand side, -1 right-hand side
    int q;     // derivative order the coefficient multiplies
    ld h;      // normalised offset from the scheme's point
};

ld fact(int n)
{
    ld r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

// Solve the Taylor-moment system of a compact scheme:
//   sum_u side_u c_u h_u^(p-q_u)/(p-q_u)! = 0 for p = 0..pmax, plus the fixed
// entries; Gaussian elimination with partial pivoting in long double.
bool solve_scheme(const std::vector<Unknown> &u, const std::vector<std::pair<int, ld>> &fixed,
                  int pmax, double *out)
{
    const int n = static_cast<int>(u.size());
    if (pmax + 1 + static_cast<int>(fixed.size()) != n) return false;
    std::vector<ld> a(n * n, 0), b(n, 0);
    int r = 0;
    for (int p = 0; p <= pmax; ++p, ++r)
        for (int c = 0; c < n; ++c)
            if (p >= u[c].q) {
                int e = p - u[c].q;
                ld hp = 1;
                for (int t = 0; t < e; ++t) hp *= u[c].h;
                a[r * n + c] = static_cast<ld>(u[c].side) * hp / fact(e);
            }
    for (auto &f : fixed) {
        a[r * n + f.first] = 1;
        b[r] = f.second;
        ++r;
    }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int i = c + 1; i < n; ++i)
            if (fabsl(a[i * n + c]) > fabsl(a[piv * n + c])) piv = i;
        if (a[piv * n + c] == 0) return false;
        if (piv != c) {
            for (int k = 0; k < n; ++k) std::swap(a[c * n + k], a[piv * n + k]);
            std::swap(b[c], b[piv]);
        }
        for (int i = c + 1; i < n; ++i) {
            ld f = a[i * n + c] / a[c * n + c];
            if (f == 0) continue;
            for (int k = c; k < n; ++k) a[i * n + k] -= f * a[c * n + k];
            b[i] -= f * b[c];
        }
    }
    std::vector<ld> x(n);
    for (int i = n - 1; i >= 0; --i) {
        ld s = b[i];
        for (int k = i + 1; k < n; ++k) s -= a[i * n + k] * x[k];
        x[i] = s / a[i * n + i];
    }
    for (int i = 0; i < n; ++i) {
        // entries below 1e-14 in magnitude are zeroed (solvers.f90:126, 189)
        out[i] = (fabsl(x[i]) < 1e-14L) ? 0.0 : static_cast<double>(x[i]);
    }
    return true;
}

// positions reached by walking interval lengths s0, s1, ...
std::vector<ld> walk(std::initializer_list<ld> steps, int sign = +1)
{
    std::vector<ld> out;
    ld acc = 0;
    for (ld s : steps) {
        acc += s;
        out.push_back(sign * acc);
    }
    return out;
}

void add_f(std::vector<Unknown> &u, const std::vector<ld> &pts)
{
    for (ld h : pts) u.push_back({-1, KF, h});
}

}  // namespace

// Scheme definitions: derivative.f90:15-1613, 2002-2081.  The normalised
// interval used by each row's metric scaling has length 1; successive
// intervals grow by the stretch s.
bool generate_coefs(double stf, int levels, tbl_coefs *c)
{
    std::memset(c, 0, sizeof(*c));
    const ld s = stf;
    bool ok = true;
    {   // first derivative, first point, explicit: 12 f'_1 = sum d f(1..5)
        std::vector<Unknown> u{{+1, KD1, 0}};
        std::vector<ld> pts{0};
        for (ld h : walk({1, s, s * s, s * s * s})) pts.push_back(h);
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 12}}, 4, c->fp_fd_e);
    }
    {   // second point: f'_1 + b f'_2 + c f'_3 = sum d f(1..6), d(f_2) = -2.5
        auto up = walk({s, s * s, s * s * s, s * s * s * s});
        std::vector<Unknown> u{{+1, KD1, -1}, {+1, KD1, 0}, {+1, KD1, up[0]}};
        std::vector<ld> pts{-1, 0};
        pts.insert(pts.end(), up.begin(), up.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}, {4, -2.5L}}, 6, c->sp_fd);
    }
    const ld cm2 = -(1 + s), cm1 = -s, cp1 = s * s, cp2 = s * s + s * s * s;
    {   // centre point, b = 1
        std::vector<Unknown> u{{+1, KD1, cm1}, {+1, KD1, 0}, {+1, KD1, cp1}};
        add_f(u, {cm2, cm1, 0, cp1, cp2});
        ok &= solve_scheme(u, {{1, 1}}, 6, c->cp_fd);
    }
    {   // penultimate point jmax-1: coefficient of f'_jmax = 1, d(f_jmax-1) = 2.5
        auto dn = walk({s * s * s, s * s, s, 1}, -1);
        const ld top = s * s * s * s;
        std::vector<Unknown> u{{+1, KD1, top}, {+1, KD1, 0}, {+1, KD1, dn[0]}};
        std::vector<ld> pts{top, 0};
        pts.insert(pts.end(), dn.begin(), dn.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}, {4, 2.5L}}, 6, c->pp_fd);
    }
    {   // last point jmax
        auto dn = walk({s * s * s, s * s, s, 1}, -1);
        std::vector<Unknown> u{{+1, KD1, 0}, {+1, KD1, dn[0]}};
        std::vector<ld> pts{0};
        pts.insert(pts.end(), dn.begin(), dn.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}}, 5, c->lp_fd);
    }
    {   // second derivative, first point
        std::vector<ld> pts{0};
        for (ld h : walk({1, s, s * s, s * s * s, s * s * s * s})) pts.push_back(h);
        std::vector<Unknown> u{{+1, KD2, 0}, {+1, KD2, pts[1]}};
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}}, 6, c->fp_sd);
    }
    {   // second point, d(f_3) = 5.25
        auto up = walk({s, s * s, s * s * s, s * s * s * s, s * s * s * s * s});
        std::vector<Unknown> u{{+1, KD2, -1}, {+1, KD2, 0}, {+1, KD2, up[0]}};
        std::vector<ld> pts{-1, 0};
        pts.insert(pts.end(), up.begin(), up.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}, {5, 5.25L}}, 7, c->sp_sd);
    }
    {   // centre point
        std::vector<Unknown> u{{+1, KD2, cm1}, {+1, KD2, 0}, {+1, KD2, cp1}};
        add_f(u, {cm2, cm1, 0, cp1, cp2});
        ok &= solve_scheme(u, {{1, 1}}, 6, c->cp_sd);
    }
    {   // penultimate point, d(f_jmax-2) = 5.25
        auto dn = walk({s * s * s * s, s * s * s, s * s, s, 1}, -1);
        const ld top = s * s * s * s * s;
        std::vector<Unknown> u{{+1, KD2, top}, {+1, KD2, 0}, {+1, KD2, dn[0]}};
        std::vector<ld> pts{top, 0};
        pts.insert(pts.end(), dn.begin(), dn.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}, {5, 5.25L}}, 7, c->pp_sd);
    }
    {   // last point
        auto dn = walk({s * s * s * s, s * s * s, s * s, s, 1}, -1);
        std::vector<Unknown> u{{+1, KD2, 0}, {+1, KD2, dn[0]}};
        std::vector<ld> pts{0};
        pts.insert(pts.end(), dn.begin(), dn.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}}, 6, c->lp_sd);
    }
    // wall tables (derivative.f90:1618-1757, 1887-1998): rows 1..7 from the wall
    const auto hp = walk({1, s, s * s, s * s * s, s * s * s * s, s * s * s * s * s});
    {   // dwydy: f'(wall) = sum_j c(j+1) f(j), c(1) = 1
        std::vector<Unknown> u{{+1, KD1, 0}};
        std::vector<ld> pts{0};
        pts.insert(pts.end(), hp.begin(), hp.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{0, 1}}, 6, c->dwydy);
    }
    {   // w_at_w: c(1) f''(wall) + c(2) f'(wall) = -sum_j c(j+2) f(j), c(9) = -1/18.
        // The reference's first-moment row also carries -c(3), which moves
        // only c(2); c(2) enters no formula of the path.
        std::vector<Unknown> u{{+1, KD2, 0}, {+1, KD1, 0}};
        std::vector<ld> pts{0};
        pts.insert(pts.end(), hp.begin(), hp.end());
        add_f(u, pts);
        ok &= solve_scheme(u, {{8, -1.0L / 18.0L}}, 7, c->w_at_w);
        ld c2 = c->w_at_w[2];
        for (int n = 0; n < 6; ++n) c2 += static_cast<ld>(c->w_at_w[3 + n]) * hp[n];
        c->w_at_w[1] = (fabsl(c2) < 1e-14L) ? 0.0 : static_cast<double>(c2);
    }
    // Poisson schemes per level at stretch stf**(2**(l-1)) (coef_generation.f90:32-38)
    for (int l = 1; l <= levels && l <= TBL_MAX_LEVELS; ++l) {
        double sd = 1.0;
        for (int t = 0; t < (1 << (l - 1)); ++t) sd *= stf;
        const ld S = sd;
        {   // row 2: a f'_1 + (4/5) f''_2 + c f''_3 = sum d f(1..4)
            std::vector<Unknown> u{{+1, KD1, -1}, {+1, KD2, 0}, {+1, KD2, S}};
            add_f(u, {-1, 0, S, S + S * S});
            ok &= solve_scheme(u, {{1, 4.0L / 5.0L}}, 5, c->sp_poi[l - 1]);
        }
        {   // centre rows, B-weight 11/15 on the diagonal
            std::vector<Unknown> u{{+1, KD2, -S}, {+1, KD2, 0}, {+1, KD2, S * S}};
            add_f(u, {-(1 + S), -S, 0, S * S, S * S + S * S * S});
            ok &= solve_scheme(u, {{1, 11.0L / 15.0L}}, 6, c->cp_poi[l - 1]);
        }
        {   // row ny-1: three-point second derivative
            std::vector<Unknown> u{{+1, KD2, 0}};
            add_f(u, {-1, 0, S});
            ok &= solve_scheme(u, {{0, 1}}, 2, c->pp_poi[l - 1]);
        }
        {   // row ny: f'' = l3 f + l4 f(-S) + l5 f(-S-1) + l2 f'
            std::vector<Unknown> u{{+1, KD2, 0}, {-1, KD1, 0}};
            add_f(u, {0, -S, -(S + 1)});
            ok &= solve_scheme(u, {{0, 1}}, 3, c->lp_poi[l - 1]);
        }
    }
    return ok;
}

static double ipow(double x, int n)
{
    double r = 1.0, b = x;
    while (n > 0) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return r;
}

// create_ctes (main_par.f:1828-1892)
void build_metrics(const tbl_grid &g, Metrics &m)
{
    m.levels = g.levels;
    m.nx.assign(g.levels, 0);
    m.ny.assign(g.levels, 0);
    m.dy0.assign(g.levels, 0);
    m.stf.assign(g.levels, 0);
    m.dx2.assign(g.levels, 0);
    m.qdy.assign(g.ny, 0.0);
    m.qdy2.assign(static_cast<size_t>(g.levels) * g.ny, 0.0);
    for (int l = 1; l <= g.levels; ++l) {
        int e = 1 << (l - 1);
        m.dy0[l - 1] = (g.stf != 1.0) ? g.dy0 * ((ipow(g.stf, e) - 1.0) / (g.stf - 1.0))
                                      : g.dy0 * static_cast<double>(e);
        m.stf[l - 1] = ipow(g.stf, e);
        double t = g.dx * static_cast<double>(e);
        m.dx2[l - 1] = t * t;
    }
    for (int j = 1; j <= g.ny - 1; ++j)
        m.qdy[j - 1] = (g.stf != 1.0) ? 1.0 / (m.dy0[0] * ipow(m.stf[0], j - 1))
                                      : 1.0 / m.dy0[0];
    for (int l = 1; l <= g.levels; ++l) {
        int nj = (g.ny - 1) / (1 << (l - 1));
        for (int j = 1; j <= nj; ++j) {
            double aux = (g.stf != 1.0) ? m.dy0[l - 1] * ipow(m.stf[l - 1], j - 1)
                                        : m.dy0[l - 1];
            m.qdy2[static_cast<size_t>(l - 1) * g.ny + (j - 1)] = 1.0 / (aux * aux);
        }
    }
    m.nx[0] = g.nx;
    m.ny[0] = g.ny;
    for (int l = 1; l < g.levels; ++l) {
        m.nx[l] = (m.nx[l - 1] + 1) / 2;
        m.ny[l] = (m.ny[l - 1] + 1) / 2;
    }
}

// verifica (main_par.f:60-134) at one rank, plus the limits of this
// implementation's kernels.
int check_grid(const tbl_grid &g)
{
    if (g.levels < 2 || g.levels > TBL_MAX_LEVELS) return TBL_ERR_GRID;
    if (g.nx < 7 || g.ny < 8 || g.nmodes < 1) return TBL_ERR_GRID;
    if (!(g.dx > 0) || !(g.dy0 > 0) || !(g.stf > 0) || !(g.fac_y > 0)) return TBL_ERR_GRID;
    int d = 1 << (g.levels - 1);
    if ((g.nx - 1) % d || (g.ny - 1) % d) return TBL_ERR_GRID;
    if (g.ny < (1 << g.levels)) return TBL_ERR_GRID;
    // coarsest level: the border phase needs columns 2 and nx-1 to be
    // independent and rows 1..5 to exist
    if ((g.nx - 1) / d + 1 < 8 || (g.ny - 1) / d + 1 < 5) return TBL_ERR_GRID;
    if (g.ny > kMaxNy) return TBL_ERR_GRID;
    // kernels address up to two modes' planes of a level through one buffer
    // resource (32-bit byte offsets, 0x80000000 the out-of-range offset):
    // keep four finest-level FP64 planes below 2 GiB
    const long long plane = static_cast<long long>(g.nx) * (((g.ny + 7) / 8) * 8);
    if (4 * plane * 16 >= (1LL << 31)) return TBL_ERR_GRID;
    return TBL_OK;
}

// coef_center / coef_border (poisson_par.f:229-340) followed by the banded
// LU without pivoting of band5_poi (penta_par.f:308-364).  Output per row j:
// l1 = multiplier of row j-1, l2 = multiplier of row j-2 (forward
// substitution y_j = r_j - l1_j y_{j-1} - l2_j y_{j-2}); u1, u2 = the two
// super-diagonals of U; piv = the pivot (U's diagonal).
void line_lu(const tbl_grid &g, const tbl_coefs &c, const Metrics &m, int lvl, int k,
             bool border, double *l1, double *l2, double *u1, double *u2, double *piv)
{
    const int L = lvl - 1, n = m.ny[L];
    const double *sp = c.sp_poi[L], *cp = c.cp_poi[L], *pp = c.pp_poi[L], *lp = c.lp_poi[L];
    const double *q2 = &m.qdy2[static_cast<size_t>(L) * g.ny];
    const double k2b2 = -static_cast<double>(k) * static_cast<double>(k) * g.beta * g.beta;
    const double qx = (border ? -15.0 : -30.0) / (12.0 * m.dx2[L]);
    const double add = (qx + k2b2) / g.fac_y;
    // band rows: a[j][0..4] = coefficients of v(j-2..j+2), j 0-based
    std::vector<double> a(static_cast<size_t>(n) * 5, 0.0);
    auto A = [&](int j1, int c1) -> double & { return a[(j1 - 1) * 5 + (c1 - 1)]; };
    A(1, 3) = 1.0;
    A(2, 2) = sp[3] * q2[0];
    A(2, 3) = sp[4] * q2[0] + sp[1] * add;
    A(2, 4) = sp[5] * q2[0] + sp[2] * add;
    A(2, 5) = sp[6] * q2[0];
    for (int j = 3; j <= n - 2; ++j) {
        A(j, 1) = cp[3] * q2[j - 3];
        A(j, 2) = cp[4] * q2[j - 3] + cp[0] * add;
        A(j, 3) = cp[5] * q2[j - 3] + cp[1] * add;
        A(j, 4) = cp[6] * q2[j - 3] + cp[2] * add;
        A(j, 5) = cp[7] * q2[j - 3];
    }
    A(n - 1, 2) = pp[1] * q2[n - 3];
    A(n - 1, 3) = pp[2] * q2[n - 3] + add;
    A(n - 1, 4) = pp[3] * q2[n - 3];
    A(n, 1) = lp[4] * q2[n - 3];
    A(n, 2) = lp[3] * q2[n - 3];
    A(n, 3) = lp[2] * q2[n - 3] + add
              - lp[1] * std::sqrt(q2[n - 3]) * std::sqrt((g.alpha * g.alpha - k2b2) / g.fac_y);
    // elimination in band storage: row j keeps (sub2, sub1, diag, sup1, sup2)
    const double tiny = static_cast<double>(1.e-20f);  // penta_par.f:317
    for (int j = 0; j < n; ++j) {
        l1[j] = l2[j] = u1[j] = u2[j] = 0.0;
    }
    // work on a dense copy of the 5 diagonals
    for (int kk = 1; kk <= n; ++kk) {
        double pk = A(kk, 3);
        if (pk == 0.0) pk = A(kk, 3) = tiny;
        for (int i = kk + 1; i <= kk + 2 && i <= n; ++i) {
            int off = i - kk;                       // 1 or 2
            double &sub = A(i, 3 - off);            // coefficient of v(kk) in row i
            double mult = sub / pk;
            if (off == 1) l1[i - 1] = mult; else l2[i - 1] = mult;
            // row_i -= mult * row_kk over the columns kk..kk+2
            for (int col = 0; col <= 2; ++col) {
                int ci = 3 - off + col;             // position of v(kk+col) in row i
                if (ci > 5) break;
                double rk = A(kk, 3 + col);
                A(i, ci) = A(i, ci) - mult * rk;
            }
            sub = 0.0;
        }
        piv[kk - 1] = A(kk, 3);
        u1[kk - 1] = (kk + 1 <= n) ? A(kk, 4) : 0.0;
        u2[kk - 1] = (kk + 2 <= n) ? A(kk, 5) : 0.0;
    }
}

// Thomas factors of a constant tridiagonal LHS (tridy 947-975): the forward
// sweep's bet_j and gam_j depend only on the matrix.
void thomas_factors(const std::vector<double> &a, const std::vector<double> &b,
                    const std::vector<double> &c, std::vector<double> &gam,
                    std::vector<double> &bet)
{
    const int n = static_cast<int>(b.size());
    gam.assign(n, 0.0);
    bet.assign(n, 0.0);
    double bt = b[0];
    bet[0] = bt;
    for (int j = 1; j < n; ++j) {
        gam[j] = c[j - 1] / bt;
        bt = b[j] - a[j] * gam[j];
        bet[j] = bt;
    }
}

// coefx / coeffx (derivs_par.f:1145-1210), coefy / coefy_fv / coeffy
// (derivs_par.f:1040-1142).  kind: 0 x first, 1 x second, 2 y first,
// 3 y second, 4 y first with the wall row forced (deryfv).
void deriv_lhs(const tbl_grid &g, const tbl_coefs &c, int kind, std::vector<double> &a,
               std::vector<double> &b, std::vector<double> &cc)
{
    const int n = (kind <= 1) ? g.nx : g.ny;
    a.assign(n, 0);
    b.assign(n, 0);
    cc.assign(n, 0);
    if (kind == 0 || kind == 1) {
        const bool sec = (kind == 1);
        for (int i = 0; i < n; ++i) {
            a[i] = sec ? 2 : 1;
            b[i] = sec ? 11 : 3;
            cc[i] = sec ? 2 : 1;
        }
        if (sec) {
            a[0] = 0; b[0] = 13; cc[0] = 137;
            a[1] = 1; b[1] = 12; cc[1] = 3;
            a[n - 2] = 3; b[n - 2] = 12; cc[n - 2] = 1;
            a[n - 1] = 137; b[n - 1] = 13; cc[n - 1] = 0;
        } else {
            a[0] = 0; b[0] = 1; cc[0] = 0;
            a[1] = 1; b[1] = 6; cc[1] = 2;
            a[n - 2] = 2; b[n - 2] = 6; cc[n - 2] = 1;
            a[n - 1] = 4; b[n - 1] = 1; cc[n - 1] = 0;
        }
        return;
    }
    const bool sec = (kind == 3);
    const double *sp = sec ? c.sp_sd : c.sp_fd, *cp = sec ? c.cp_sd : c.cp_fd;
    const double *pp = sec ? c.pp_sd : c.pp_fd, *lp = sec ? c.lp_sd : c.lp_fd;
    for (int j = 0; j < n; ++j) {
        a[j] = cp[0];
        b[j] = cp[1];
        cc[j] = cp[2];
    }
    if (kind == 2) { a[0] = 0; b[0] = c.fp_fd_e[0]; cc[0] = 0; }
    else if (kind == 4) { a[0] = 0; b[0] = 1.0; cc[0] = 0; }
    else { a[0] = 0; b[0] = c.fp_sd[0]; cc[0] = c.fp_sd[1]; }
    a[1] = sp[0]; b[1] = sp[1]; cc[1] = sp[2];
    a[n - 2] = pp[2]; b[n - 2] = pp[1]; cc[n - 2] = pp[0];
    a[n - 1] = lp[1]; b[n - 1] = lp[0]; cc[n - 1] = 0;
}

}  // namespace tbl

// ---------------------------------------------------------------------------
// coefs.bin reader: the 20 sequential unformatted records written by
// pre_processing/coef_generation.f90:39-60 and read by main_par.f:852-873
// (4-byte record markers, little endian).  Records 1-14 are the tables of
// the hot path; records 15-20 (wall vorticity, integration) are skipped.
// ---------------------------------------------------------------------------
#include <cstdio>

namespace tbl {

int read_coefs_bin(const char *path, int levels, tbl_coefs *c, std::string *err)
{
    std::FILE *f = std::fopen(path, "rb");
    if (!f) {
        *err = "cannot open coefs file";
        return TBL_ERR_ARG;
    }
    std::vector<std::vector<double>> rec;
    for (;;) {
        int32_t n1 = 0, n2 = 0;
        if (std::fread(&n1, 4, 1, f) != 1) break;
        if (n1 < 0 || n1 % 8) {
            std::fclose(f);
            *err = "bad record marker";
            return TBL_ERR_ARG;
        }
        std::vector<double> r(n1 / 8);
        if (std::fread(r.data(), 1, n1, f) != static_cast<size_t>(n1) ||
            std::fread(&n2, 4, 1, f) != 1 || n2 != n1) {
            std::fclose(f);
            *err = "truncated record";
            return TBL_ERR_ARG;
        }
        rec.push_back(std::move(r));
    }
    std::fclose(f);
    if (rec.size() < 14) {
        *err = "fewer than 14 records";
        return TBL_ERR_ARG;
    }
    if (rec[0].size() != 6) {
        *err = "record 1 must hold the 6-entry explicit first-point scheme (a 7-entry record "
               "is the older implicit scheme)";
        return TBL_ERR_ARG;
    }
    const size_t want[10] = {6, 9, 8, 9, 7, 8, 10, 8, 10, 8};
    double *dst[10] = {c->fp_fd_e, c->sp_fd, c->cp_fd, c->pp_fd, c->lp_fd,
                       c->fp_sd, c->sp_sd, c->cp_sd, c->pp_sd, c->lp_sd};
    std::memset(c, 0, sizeof(*c));
    for (int i = 0; i < 10; ++i) {
        if (rec[i].size() != want[i]) {
            *err = "derivative record " + std::to_string(i + 1) + " has the wrong length";
            return TBL_ERR_ARG;
        }
        std::memcpy(dst[i], rec[i].data(), want[i] * sizeof(double));
    }
    if (rec.size() >= 16) {     // the wall tables (coef_generation.f90:54-55)
        if (rec[14].size() != 9 || rec[15].size() != 8) {
            *err = "wall records 15-16 must hold 9 and 8 entries";
            return TBL_ERR_ARG;
        }
        std::memcpy(c->w_at_w, rec[14].data(), 9 * sizeof(double));
        std::memcpy(c->dwydy, rec[15].data(), 8 * sizeof(double));
    }
    const int per[4] = {7, 8, 4, 5};
    for (int t = 0; t < 4; ++t) {
        const auto &r = rec[10 + t];
        if (r.size() != static_cast<size_t>(per[t]) * levels) {
            *err = "Poisson record " + std::to_string(11 + t) + " does not hold `levels` levels";
            return TBL_ERR_ARG;
        }
        for (int l = 0; l < levels; ++l)
            for (int e = 0; e < per[t]; ++e) {
                const double v = r[static_cast<size_t>(l) * per[t] + e];
                if (t == 0) c->sp_poi[l][e] = v;
                else if (t == 1) c->cp_poi[l][e] = v;
                else if (t == 2) c->pp_poi[l][e] = v;
                else c->lp_poi[l][e] = v;
            }
    }
    return TBL_OK;
}

}  // namespace tbl
