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
    // Kelvin-Voigt coating: core radii, compression (current, previous step),
    // coating friction history z (component-major [2][cap]), stick, |q_trial|,
    // dissipation; cneed: compression > 0 seen by the probe (pattern input)
    double* cra; double* crb;
    double* cc; double* ccp;
    double* cz;
    int8_t* cstick;
    double* cqtn; double* cdiss;
    int8_t* cneed;
    int64_t cap;
};

struct PairParams {
    double k_pen = 0.0, delta_h = 0.0, k_t_ratio = 0.0, mu_eff = 0.0, g_on = 0.0, g_off = 0.0;
    int allow_deact = 0;
    int n_sinc = 0;
    // per-element friction coefficient (nullable): pair i uses mu_elem[ea[i]]
    // scaled by mu_eff instead of mu_eff (batches of independent solves)
    const double* mu_elem = nullptr;
    const int* ea = nullptr;
    // Kelvin-Voigt coating (coat_on = 0: off): stiffness, damping, friction
    // coefficient, tangential stiffness k * k_t_ratio, dashpot time step
    int coat_on = 0;
    double coat_k = 0.0, coat_c = 0.0, coat_mu = 0.0, coat_kt = 0.0, coat_dt = 0.0;
    int* flags = nullptr;   // status flags (FLAG_COAT_UNPLACED)
    // position of (slot s, node k) in the node contribution lists: nqpos[4 s + k]
    const int* nqpos = nullptr;
    // device Newton control state word (nullable): the kernel does nothing unless it is 0
    const double* stop = nullptr;
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

// Arguments of the one-launch residual / tangent gather (compose.hip).
struct GatherArgs {
    int n_nodes, n_elems, n_sinc, n_blocks;
    int node_mode;     // -1 none, 0 f_int, 1 pair force, 2 full residual
    int node_blocks;   // set by the launcher
    int block;         // workgroup size of the gather (128, 256 or 512)
    NodeLists L;
    BlockLists B;
    const double* fe;  // element force staging (general path)
    const double* fint;  // element-kernel internal force (direct path), replaces fe
    // direct path: descriptors of the blocks with pair contributions,
    // {CSR offset of the block's first entry, scalar row stride, psrc begin,
    // psrc end | bit 31 if the element kernel wrote a beam value there}
    int direct;
    const int4* cdesc;
    int n_cdesc;
    const double* nf;  // pair force staging: c_k F (3 values) per node-list entry
    const double* u;
    const double* f_ext;
    const uint8_t* fixed;
    DynArgs dy;
    double* out;       // residual / force output (6 n_nodes)
    int* flags;
    int use_beam, use_pair;
    double mass_coef;
    const double* M;
    // norm / active-count reduction (partial == nullptr: none)
    double* partial;       // node_blocks partial sums of R^2
    double* res_norm2;
    const int* act_part;   // per-workgroup active counts of the pair kernel
    int n_act_part;
    int* n_active;
    const double* stop;    // device Newton control state word (nullable): skip once the call ended
};

}  // namespace xk
