// Host-side setup routines of trans_bl_amd (see tbl_host.cpp).
#pragma once

#include <string>
#include <vector>

#include "../../include/trans_bl_amd.h"

namespace tbl {

constexpr int kMaxNy = 1024;   // largest jmax the line kernels accept

struct Metrics {
    int levels = 0;
    std::vector<int> nx, ny;                  // v_ptsx, v_ptsy
    std::vector<double> dy0, stf, dx2;        // v_dy0, v_stf, v_dx2
    std::vector<double> qdy;                  // v_qdy(j)
    std::vector<double> qdy2;                 // v_qdy2(j, l) at [l*ny + j]
};

bool generate_coefs(double stf, int levels, tbl_coefs *c);
void build_metrics(const tbl_grid &g, Metrics &m);
int check_grid(const tbl_grid &g);
void line_lu(const tbl_grid &g, const tbl_coefs &c, const Metrics &m, int lvl, int k,
             bool border, double *l1, double *l2, double *u1, double *u2, double *piv);
void thomas_factors(const std::vector<double> &a, const std::vector<double> &b,
                    const std::vector<double> &c, std::vector<double> &gam,
                    std::vector<double> &bet);
int read_coefs_bin(const char *path, int levels, tbl_coefs *c, std::string *err);
void deriv_lhs(const tbl_grid &g, const tbl_coefs &c, int kind, std::vector<double> &a,
               std::vector<double> &b, std::vector<double> &cc);

}  // namespace tbl
