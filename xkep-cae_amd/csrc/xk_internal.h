// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <string>
#include <vector>

#include "xk_common.h"

namespace xk {

void set_error(const std::string& s);

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t nbytes);
    void release();
};

struct PairSoA;
struct PairParams;
struct NodeLists;
struct BlockLists;
struct DynArgs;

}  // namespace xk

// complete definitions needed by Ctx
#include "xk_structs.h"

namespace xk {

enum PairBuf {
    PB_NA0, PB_NA1, PB_NB0, PB_NB1, PB_RA, PB_RB, PB_S, PB_T, PB_GAP, PB_N, PB_T1,
    PB_PN, PB_KPEN, PB_KT, PB_QTN, PB_DISS, PB_Z, PB_STATUS, PB_STICK, PB_SINC, PB_EA, PB_EB,
    PB_CRA, PB_CRB, PB_CC, PB_CCP, PB_CZ, PB_CSTICK, PB_CQTN, PB_CDISS, PB_CNEED,
    PB_COUNT
};

constexpr int PROF_EV = 5;   // events per profiled assembly call: 4 stages

struct Ctx {
    int n_nodes = 0, n_elems = 0, prec = 0;
    std::vector<double> X_h, rad_h;
    std::vector<int> conn_h;
    BeamConst bc{};
    DBuf d_X, d_conn, d_rad, d_node_bptr, d_node_bsrc, d_fe, d_Ke, d_flags, d_partial;
    // pairs
    std::vector<DBuf> pb;
    PairSoA P{};
    int* ea = nullptr;
    int* eb = nullptr;
    int64_t n_pairs = 0, cap = 0;
    int last_modes = 0;
    // tangent pattern of the current increment
    bool pattern_valid = false;
    int n_sinc = 0, n_blocks = 0;
    std::vector<int> gptr_h, gcol_h;   // host copy, valid when host_pattern_valid
    bool host_pattern_valid = false;
    bool host_build = false;   // direct path: build the pattern on the host (XK_OPT_HOST_PATTERN)
    DBuf d_nflag;          // direct path, per node: bit0 has an element, bit1 an element joins n and n + 1
    DBuf d_bbeam;          // direct path, per block: 1 if the beam part holds it
    std::vector<DBuf> pat;  // scratch of the device pattern build
    DBuf pat_tmp;           // radix-sort scratch
    DBuf d_gptr, d_gcol, d_brow, d_blk_bptr, d_blk_bsrc, d_blk_pptr, d_blk_psrc, d_node_pptr, d_node_psrc;
    DBuf d_nf, d_sg;
    DBuf d_node_qpos;   // position of (slot, k) in the node lists (inverse of d_node_psrc)
    // broadphase scratch and candidate bookkeeping
    DBuf bp[18];
    DBuf gj;                        // constraint Jacobian: per-pair flags, row lengths, offsets
    int64_t last_cand = 0;
    int64_t raw_n = 0, raw_cap = 1;
    double last_cell = 0.0;
    double cell_factor = 1.5;       // grid cell = factor x mean box extent
    bool filtered_valid = false;    // filtered list of the last xk_prescreen_radii still held
    bool keys_valid = false;        // bp[14] holds the sorted keys of the pair list
    const int* cand_I = nullptr;
    const int* cand_J = nullptr;
    // optional per-stage event timing of the fused assembly
    std::vector<hipEvent_t> prof_ev;   // PROF_EV per profiled call
    int prof_max = 0, prof_calls = 0;
    // second stream (default priority) for the element kernel, joined by events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DBuf d_act;   // per-workgroup active counts of the pair kernel
    // direct path (chain-structured meshes): element kernel writes K / f_int itself
    bool direct = false;
    DBuf d_eflag;   // per element: bit0 shares node1 with element e-1, bit1 shares node2 with e+1
    DBuf d_eslot;   // per element: global block index of its 16 group blocks
    DBuf d_cdesc;   // int4 descriptors of the blocks with pair contributions
    DBuf d_fint;    // internal force written by the element kernel (6 n_nodes)
    int n_cdesc = 0;
    // element chains (direct path): maximal runs of nodes joined n -> n + 1
    int n_chains = 0;
    std::vector<int> cstart_h, node_chain_h;
    DBuf d_cstart, d_node_chain;
    // device linear solve: per-pattern set-up, buffers, dense capacitance blocks
    bool solve_ready = false;
    bool tangent_staged = false;   // the last fused assembly staged the pairs' tangent records on this pattern
    std::vector<DBuf> sv;
    std::vector<int> sv_comp_n;                    // contributing pairs per component
    std::vector<int64_t> sv_comp_pos, sv_comp_soff;  // first position / capacitance offset per component
    std::vector<char> sv_comp_piv;                  // per component: its LU pivots rows (partial pivoting)
    void* sv_handle = nullptr;                     // rocblas handle
    const uint8_t* sv_fixed = nullptr;
    std::vector<int64_t> piv_h;   // host copies of the capacitance pivots and their permutation
    std::vector<int> perm_h;
    std::vector<int64_t> info_h;
    int sv_lu_kind = 0;        // capacitance LU: 0 without interchanges (pivoting on failure), 1 partial pivoting
    bool sv_gepp = false;      // this pattern's solves use partial pivoting (after a failure of the other)
    int64_t sv_fallbacks = 0;  // solves that fell back to partial pivoting
    // per-element friction coefficients (xk_ctx_set_friction)
    DBuf d_mu_elem;
    bool has_mu_elem = false;
    // Kelvin-Voigt coating (xk_ctx_set_coating): per-element core radii, parameters,
    // the dashpot time step of the next assemblies (xk_coating_set_dt)
    std::vector<double> crad_h;
    DBuf d_crad;
    DBuf d_coat_cnt;   // probe: #(pairs with compression > 0 and no slot)
    bool coat_on = false;
    double coat_k = 0.0, coat_c = 0.0, coat_mu = 0.0, coat_ktr = 0.5, coat_dt = 1e-15;
    void coat_params(PairParams& pp) const {
        pp.coat_on = coat_on ? 1 : 0;
        pp.coat_k = coat_k;
        pp.coat_c = coat_c;
        pp.coat_mu = coat_mu;
        pp.coat_kt = coat_k * coat_ktr;
        pp.coat_dt = coat_dt;
        pp.flags = (int*)d_flags.p;
    }
    // launch shapes (xk_ctx_set_tuning)
    int pair_block = 256, gather_block = 256;
    // learned prescreen: parameters, features, watch / prune lists of the increment
    DBuf lp_prm;        // FP32 parameters
    float lp_thr = 0.f;
    bool lp_ready = false;
    DBuf lp_feat, lp_gprev, lp_keep, lp_pos, lp_watch, lp_prune, lp_count;
    // translational displacement of every assembly over the watch list within a
    // Newton call (LP_HIST_CAP slots of 3 n_nodes doubles) and the slot counter
    DBuf lp_hist, lp_hist_cnt;
    static constexpr int LP_HIST_CAP = 64;
    int64_t lp_gprev_n = 0, lp_gprev_cap = 0;
    int64_t n_watch = 0, n_prune = 0;
    bool lp_on = false;          // the fused assembly runs the pair kernel over the watch list
    // while lp_on, the watched pairs' state lives in these compact arrays (pair
    // order of the watch list, capacity n_watch) and PW points at them
    std::vector<DBuf> wb;
    PairSoA PW{};
    // pair-state snapshot (xk_pairs_snapshot / xk_pairs_restore)
    std::vector<DBuf> snap;
    int64_t snap_n = -1;

    int grow_pairs(int64_t need);
    int lp_compact(hipStream_t st);                 // gather the watched pairs into wb / PW
    int lp_sync(hipStream_t st, bool keep_on);      // scatter them back (and end the watch list)
    void bind_pairs();
    int build_pattern();
    int solve_setup();
    int solve_apply(const double* r, double* x, hipStream_t st, const double* stop);
    int sv_factor_capacitance(bool gepp, hipStream_t st);
    int linear_solve(const double* u, const uint8_t* fixed, double mass_coef, const double* M, const double* R,
                     double* du, int refine, double tol, double* res_rel, int* n_refine, hipStream_t st);
    void release_solver();
    DBuf nt_try;   // trial displacement of the device line search
    DBuf nt_part;  // per-workgroup partial sums of the line search's norms
    // device Newton control record of the call in progress (xk_newton_control_attach;
    // null: none): the kernels of an attempt read its state word and skip once the call ended
    double* nstop = nullptr;
    int newton_control_begin(double* ctl, int hist_cap, double tol_force, double tol_disp, double f_ext_ref_norm,
                             int dynamic_ref, int divergence_window, int max_attempts, double solve_tol,
                             hipStream_t st);
    int newton_control_force(const double* res_norm2, const int* n_active, hipStream_t st);
    int newton_control_set(int state, int reason, hipStream_t st);
    int newton_update(double* u, double* du, const double* R, const double* f_contact, const double* f_ext,
                      const uint8_t* fixed, const double* res_norm2, int max_steps, double diverge_factor,
                      double cap, double* status, hipStream_t st);
    int solve_buffer(int which, void* host, int64_t nbytes, int64_t* size);
    int build_pattern_gpu(hipStream_t st);   // direct path: the pattern built on the device
    int host_pattern();                      // gptr_h / gcol_h of the current pattern
    int ensure_pattern();
    BlockLists block_lists() const;
    NodeLists node_lists() const;
    void release_all() {
        DBuf* all[] = {&d_X, &d_conn, &d_rad, &d_node_bptr, &d_node_bsrc, &d_fe, &d_Ke, &d_flags, &d_partial,
                       &d_gptr, &d_gcol, &d_brow, &d_blk_bptr, &d_blk_bsrc, &d_blk_pptr, &d_blk_psrc,
                       &d_node_pptr, &d_node_psrc, &d_nf, &d_sg, &d_node_qpos};
        for (DBuf* b : all) b->release();
        for (auto& b : pb) b.release();
        for (auto& b : bp) b.release();
        d_act.release();
        d_crad.release();
        d_coat_cnt.release();
        d_eflag.release();
        d_nflag.release();
        d_cstart.release();
        d_node_chain.release();
        nt_try.release();
        nt_part.release();
        release_solver();
        d_bbeam.release();
        pat_tmp.release();
        for (auto& b : pat) b.release();
        d_eslot.release();
        d_cdesc.release();
        d_fint.release();
        d_mu_elem.release();
        for (DBuf* b : {&lp_prm, &lp_feat, &lp_gprev, &lp_keep, &lp_pos, &lp_watch, &lp_prune, &lp_count}) b->release();
        for (auto& b : snap) b.release();
        for (auto& b : wb) b.release();
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        prof_ev.clear();
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        ev_fork = ev_join = nullptr;
        side = nullptr;
    }
    GatherArgs gather_args(int node_mode) const;
    // element kernel of the context's path (direct: K gets the beam blocks
    // and the mass term; staged: element records)
    // kprec: precision of K (-1: the context's; 0 FP64 for the solve's beam rows)
    hipError_t launch_elements(const double* u, bool want_f, void* K, double mass_coef, const double* M,
                               hipStream_t st, int kprec = -1);
};

// launchers (one per translation unit)
hipError_t launch_beam(int prec, int n_elems, const double* X, const double* u, const int* conn,
                       const BeamConst& bc, double* fe_stage, void* Ke_stage, int* flags, hipStream_t st,
                       const double* stop = nullptr);
// direct variant: writes f_int (nullable) and the beam blocks of K (nullable,
// pattern order, mass_coef * M added on diagonal blocks) without staging
hipError_t launch_beam_direct(int prec, int n_elems, const double* X, const double* u, const int* conn,
                              const BeamConst& bc, const int8_t* eflag, const int* eslot, const int* gptr,
                              double mass_coef, const double* M, double* f_int, void* K, int* flags,
                              hipStream_t st, const double* stop = nullptr);
// idx (nullable): the kernel visits pairs idx[0 .. n_pairs) instead of 0 .. n_pairs;
// bs: workgroup size (64, 128, 256 or 512), act_part holds pair_blocks(n, bs) counts
hipError_t launch_pairs(int prec, int64_t n_pairs, const PairSoA& P, const double* Xr, const double* u,
                        const double* u_ref, int modes, const PairParams& pp, double* nf, void* sg,
                        int* act_part, hipStream_t st, const int* idx = nullptr, int bs = 256);
bool pair_block_ok(int bs);
bool gather_block_ok(int bs);
// exclusive scan of n int64 values (scratch slots bp[s0..]); total to *total_host (synchronous)
int scan_exclusive_i64(Ctx* c, int s0, const int64_t* in, int64_t* out, int64_t n, int64_t* total_host,
                       hipStream_t st);
// learned prescreen (prescreen.hip)
int lp_n_params();
int lp_grad_blocks(int64_t n);
hipError_t launch_lp_features(int64_t n, const PairSoA& P, const int* ea, const int* eb, const double* Xr,
                              const double* u, const int* elem_layer, float layer_scale, double* gprev,
                              int64_t n_prev, uint16_t* feat, float* featf, hipStream_t st);
hipError_t launch_lp_infer_bf16(int64_t n, const uint16_t* feat, const float* prm, float thr, const int* sinc,
                                int64_t* keep, float* logits, hipStream_t st);
hipError_t launch_lp_infer_f32(int64_t n, const float* feat, const float* prm, float thr, int64_t* keep,
                               float* logits, hipStream_t st);
hipError_t launch_lp_split(int64_t n, const int64_t* keep, const int64_t* pos, int* watch, int* prune,
                           hipStream_t st);
hipError_t launch_lp_miss(int64_t n, const int* idx, const double* gap, double g_on, int* count, hipStream_t st);
hipError_t launch_lp_hist_push(int n_nodes, const double* u, double* hist, int cap, int* count, hipStream_t st,
                               const double* stop = nullptr);
hipError_t launch_pair_replay(int64_t n_idx, const int* idx, const PairSoA& P, const double* Xr, const double* hist,
                              int64_t stride, int n_hist, hipStream_t st);
hipError_t launch_lp_grad(int64_t n, const float* X, const float* y, const float* prm, float pos_w, float* partial,
                          float* grad, hipStream_t st);
// arrays moved by one permute_set launch
constexpr int PERM_SET_MAX = 32;
struct PermuteSet {
    int count = 0;
    const void* a[PERM_SET_MAX];
    void* b[PERM_SET_MAX];
    int esz[PERM_SET_MAX];
    int comps[PERM_SET_MAX];
};
hipError_t launch_permute_set(const PermuteSet& s, int64_t n, int64_t sa, int64_t sb, const int* idx, int scatter,
                              hipStream_t st);
hipError_t launch_permute(int esz, int64_t n, int comps, const void* a, int64_t sa, void* b, int64_t sb,
                          const int* idx, int scatter, hipStream_t st);
hipError_t launch_lp_adam(float* prm, const float* g, float* m, float* v, int t, float lr, float b1, float b2,
                          float eps, hipStream_t st);
int pair_blocks(int64_t n_pairs, int bs);
int gather_node_blocks(int n_nodes, int bs);
// `between` (nullable) is recorded after the gather kernel, before the finish kernel
hipError_t launch_gather(int prec, GatherArgs g, const void* Ke, const void* sg, void* K, hipStream_t st,
                         hipEvent_t between = nullptr);
hipError_t launch_fill_i32(size_t n, int* p, int v, hipStream_t st);
hipError_t launch_fill_f64(size_t n, double* p, double v, hipStream_t st);
hipError_t launch_fill_f32(size_t n, float* p, float v, hipStream_t st);

}  // namespace xk
