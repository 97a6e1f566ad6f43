// Plain structures passed by value from the host launchers to the kernels.
#pragma once

#include "xk_common.h"

namespace xk {

// Pair list, structure of arrays with capacity stride `cap` for the
// component-major vectors (n: [3][cap], t1: [3][cap], z: [2][cap]).
struct PairSoA {
    int* na0; int* na1; int* nb0; int* nb1;   // node ids of segments A and B
    double* ra; double* rb;                   // radius_a, radius_b
    double* s; double* t; double* gap;
    double* n;
    double* t1;
    double* p_n; double* k_pen; double* k_t; double* qtn; double* diss;
    double* z;
    int8_t* status; int8_t* stick;
    int* sinc;                                // contributing-pair index of the increment, or -1
    int64_t cap;
};

struct PairParams {
    double k_pen, delta_h, k_t_ratio, mu_eff, g_on, g_off;
    int allow_deact;
    int n_sinc;
};

struct NodeLists {
    const int* bptr; const int* bsrc;   // element contributions: e * 2 + local node
    const int* pptr; const int* psrc;   // pair contributions: p * 4 + local node
};

struct BlockLists {
    const int* gptr;   // group-row pointer (2 n_nodes + 1)
    const int* gcol;   // block column group
    const int* brow;   // block row group
    const int* bptr; const int* bsrc;   // element contributions: e * 16 + lb
    const int* pptr; const int* psrc;   // pair contributions: p * 16 + k * 4 + l
};

struct DynArgs {
    const double* M; const double* u_pred; const double* v_pred; const double* acc_old;
    double* acc; double* vel;
    double c0, dt, gamma, alpha_m;
    int on;
};

}  // namespace xk
