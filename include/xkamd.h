/*
 * xkamd.h -- C ABI of the MI355X (gfx950) per-Newton-iteration assembly path
 * of the xkep-cae strand-contact solver.
 *
 * Plain pointers and sizes only.  Pointers documented as "device" are HIP
 * device pointers (any allocator; the Python layer passes torch tensor
 * storage).  "host" pointers are ordinary host memory.  Every call that
 * launches work takes a hipStream_t passed as void* (NULL = default stream)
 * and is asynchronous unless documented otherwise.
 *
 * Return codes: 0 = ok, > 0 = ok with a status flag (see XK_FLAG_*),
 * < 0 = error; xk_last_error() returns the message of the last error on the
 * calling thread.
 *
 * DOF layout: 6 per node [ux, uy, uz, thx, thy, thz] (the reference's
 * ndof_per_node = 6).  Element i joins conn[2i] -> conn[2i+1].
 *
 * Tangent storage: K values are written in the order of a scalar CSR matrix
 * whose pattern is a 3x3-block CSR over DOF groups (group 2n = translations
 * of node n, group 2n+1 = rotations).  xk_pattern_export() returns that
 * scalar CSR (indptr/indices) so the values form a standard CSR matrix.
 */
#ifndef XKAMD_H
#define XKAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xk_ctx xk_ctx;

/* XK_PREC_FP32: the mixed form plus the FP32 narrowphase and frame transport
 * for pairs that contribute nothing (INACTIVE, no coating), with an FP64
 * re-evaluation of every pair whose FP32 gap cannot decide the hysteresis
 * threshold or whose segments are nearly parallel; contributing pairs, forces,
 * residual and norm stay FP64.  Statuses and active sets equal the FP64 form's;
 * the stored geometry of the FP32-evaluated pairs carries FP32 rounding. */
enum { XK_PREC_FP64 = 0, XK_PREC_MIXED = 1, XK_PREC_FP32 = 2 };

/* OR-ed into xk_ctx_create's precision: use the staged element path even when
 * the connectivity is made of element chains (the direct path is chosen
 * automatically for chains: every element e joins nodes n and n + 1, every
 * node shared by two elements e and e+1 as node 2 of e and node 1 of e+1, no
 * node in more than two elements). */
enum { XK_OPT_STAGED = 256 };

/* OR-ed into xk_ctx_create's precision: on the direct path, build the tangent
 * pattern of each increment on the host instead of the device.  Both builds
 * give the same pattern and the same contribution lists (so bitwise the same
 * assembly); the option exists to check the device build against the host
 * one.  No counterpart in the reference. */
enum { XK_OPT_HOST_PATTERN = 512 };

/* status flags returned (bitwise OR) by assembly calls */
enum {
    XK_FLAG_NONFINITE = 1,     /* a non-finite residual entry was produced */
    XK_FLAG_DEGENERATE = 2,    /* an element of (near) zero length was met */
    XK_FLAG_SINGULAR = 4,      /* the device linear solve met a zero pivot */
    XK_FLAG_COAT_UNPLACED = 8  /* a pair with a coating contribution held no slot of the pattern
                                  (its terms are missing: probe, rebuild, assemble again) */
};

/* pair-kernel modes (bitwise OR) */
enum {
    XK_PAIR_GEOM = 1,          /* narrowphase + frame (UpdateGeometryProcess) */
    XK_PAIR_HYSTERESIS = 2,    /* active-set update (unfrozen UpdateGeometry) */
    XK_PAIR_HUBER = 4,         /* HuberContactForceProcess.evaluate */
    XK_PAIR_FRICTION = 8,      /* CoulombReturnMappingProcess.evaluate */
    XK_PAIR_TANGENT_C = 16,    /* stage Huber tangent blocks */
    XK_PAIR_TANGENT_F = 32,    /* stage friction tangent blocks (material + geometric) */
    /* Kelvin-Voigt coating terms (the coating on: xk_ctx_set_coating), as the
     * methods of KelvinVoigtCoatingProcess (contact/coating/strategy.py:80-384) */
    XK_PAIR_COAT_FORCE = 128,      /* .forces: k cc + c (cc - cc_prev) / dt along -n (:97-144) */
    XK_PAIR_COAT_FRICTION = 256,   /* .friction_forces: coating return mapping, history update (:206-297) */
    XK_PAIR_COAT_TANGENT = 512,    /* .stiffness: stage k_eff (w n)(w n)^T blocks (:146-204) */
    XK_PAIR_COAT_TANGENT_F = 1024  /* .friction_stiffness: stage (w w) T D T^T blocks (:299-379) */
};

/* Mesh, material and section.  Replaces ULCRBeamAssembler.__init__
 * (xkep_cae/elements/_beam_assembler.py:36-70): node_coords (n_nodes x 3,
 * host, float64), connectivity (n_elems x 2, host, int32), section
 * [E, G, A, Iy, Iz, J, kappa_y, kappa_z], scf <= 0 meaning None, v_ref NULL
 * meaning the automatic reference axis, elem_radius (n_elems, host) the
 * contact radii of DetectCandidatesInput.radii
 * (xkep_cae/contact/_manager_process.py:135-145). */
int xk_ctx_create(xk_ctx** out, int n_nodes, int n_elems, const double* node_coords,
                  const int32_t* connectivity, const double* section8, double scf,
                  const double* v_ref3, const double* elem_radius, int precision);
int xk_ctx_destroy(xk_ctx* ctx);
const char* xk_last_error(void);
int xk_ctx_precision(const xk_ctx* ctx);
/* 1 if the context uses the direct element path (element chains), 0 staged. */
int xk_ctx_direct(const xk_ctx* ctx);

/* Element assembly.  Replaces ULCRBeamAssembler.assemble_internal_force /
 * assemble_tangent (_beam_assembler.py:73-115) via assemble_cr_beam3d
 * (_beam_assembly.py:26-152): u (device, 6 n_nodes float64); f_int (device,
 * float64, nullable) receives the internal force; K_vals (device, nullable,
 * float64 or float32 by precision) receives the beam tangent in the current
 * pattern's CSR order (zeros on contact-only blocks); K_cap is its length in
 * entries and the call fails if the pattern holds more. */
int xk_beam_assemble(xk_ctx* ctx, const double* u, double* f_int, void* K_vals, int64_t K_cap,
                     void* stream);

/* Pair list (the manager.pairs of _ContactManagerInput,
 * xkep_cae/contact/_contact_pair.py:39-81), structure of arrays. */
int xk_pairs_count(const xk_ctx* ctx, int64_t* n_pairs);
/* Replace the whole pair list from host arrays (n rows each): elem_a, elem_b
 * (int32), and optionally state: s, t, gap (n), normal, tangent1 (3n),
 * p_n, k_pen, k_t, q_trial_norm, dissipation (n), z_t (2n), status, stick
 * (int8).  NULL state pointers mean the fresh _ContactStateOutput(). */
int xk_pairs_set(xk_ctx* ctx, int64_t n, const int32_t* elem_a, const int32_t* elem_b,
                 const double* s, const double* t, const double* gap, const double* normal,
                 const double* tangent1, const double* p_n, const double* k_pen,
                 const double* k_t, const double* q_trial_norm, const double* dissipation,
                 const double* z_t, const int8_t* status, const int8_t* stick);
/* Copy the pair list to host arrays (each may be NULL).  Synchronous. */
int xk_pairs_get(xk_ctx* ctx, int32_t* elem_a, int32_t* elem_b, double* s, double* t,
                 double* gap, double* normal, double* tangent1, double* p_n, double* k_pen,
                 double* k_t, double* q_trial_norm, double* dissipation, double* z_t,
                 int8_t* status, int8_t* stick);

/* DetectCandidatesProcess.process (xkep_cae/contact/_manager_process.py:172-257):
 * GPU AABB broadphase over deformed coordinates (device, n_nodes x 3,
 * float64) with inclusive-overlap output identical to _broadphase_aabb
 * (_broadphase.py:13-96), shared-node / same-layer / midpoint filters
 * (elem_layer: device int32 per element or NULL), existing pairs outside the
 * candidates set INACTIVE, new candidates appended in (i<j) order.
 * cell_size <= 0 selects the tuned cell size.  prescreen_extra is the
 * midpoint cutoff margin (the caller computes prescreening_margin or
 * 0.5 mean(r_a + r_b)); pass a negative value to disable the prescreen.
 * Synchronous (returns counts).  n_candidates receives the filtered count. */
int xk_detect_candidates(xk_ctx* ctx, const double* coords, double margin, double cell_size,
                         const int32_t* elem_layer, double prescreen_extra,
                         int64_t* n_raw, int64_t* n_candidates, void* stream);
/* The raw broadphase alone (_broadphase_aabb): pairs_out (device int32,
 * 2*cap) gets sorted (i<j) pairs.  Synchronous. */
int xk_broadphase(xk_ctx* ctx, const double* coords, double margin, double cell_size,
                  int32_t* pairs_out, int64_t cap, int64_t* n_out, void* stream);
/* The Strategy twin of candidate detection, _detect_candidates
 * (xkep_cae/contact/geometry/strategy.py:135-199, behind
 * ContactGeometryStrategy.detect, core/strategies/protocols.py:115-122):
 * replaces the whole pair list with the broadphase pairs of coords (device,
 * n_nodes x 3), filtered by shared nodes only when shared_node_filter is
 * non-zero (the Strategy's exclude_same_layer), no layer filter, no midpoint
 * prescreen, every pair with a fresh state, in sorted (i<j) order.  The
 * initial narrowphase is the caller's xk_pairs_run (GEOM | HYSTERESIS with
 * g_on = g_off = 0, _batch_update_geometry with config None, :205-305).
 * n_pairs receives the new list's length.  Synchronous. */
int xk_strategy_detect(xk_ctx* ctx, const double* coords, double margin, double cell_size,
                       int shared_node_filter, int64_t* n_pairs, void* stream);
/* Constraint Jacobian G = d g_n / d u of the pair list, replacing
 * _build_constraint_jacobian_ptp (xkep_cae/contact/geometry/strategy.py:83-132,
 * ContactGeometryStrategy.build_constraint_jacobian, protocols.py:139-147):
 * one row per pair with status != INACTIVE in pair order, entries c_k n_d at
 * DOF node_k * ndof_per_node + d of [a0, a1, b0, b1] with
 * c = [1 - s, s, -(1 - t), -t], |c_k n_d| <= 1e-30 not stored, entries on one
 * DOF summed, columns sorted (SciPy's CSR of the reference's COO).  Outputs
 * are device arrays: active_idx (n_rows, int32: the pair index of each row),
 * indptr (n_rows + 1, int64), indices (nnz, int32), data (nnz, float64); pass
 * all four NULL to query n_rows and nnz only.  ndof_total only checks that
 * the mesh's DOFs fit.  Synchronous. */
int xk_pairs_constraint_jacobian(xk_ctx* ctx, int ndof_per_node, int64_t ndof_total, int64_t* n_rows,
                                 int64_t* nnz, int32_t* active_idx, int64_t* indptr, int32_t* indices,
                                 double* data, int64_t cap_rows, int64_t cap_nnz, void* stream);
/* Candidate list of the last xk_detect_candidates (device int32 2*n). */
int xk_last_candidates(xk_ctx* ctx, int32_t* pairs_out_host, int64_t cap);
/* Sum of (r_a + r_b) over the shared-node-filtered candidates of the last
 * broadphase, in sorted (i<j) order, copied to host (n entries), for the
 * caller's mean.  Synchronous. */
int xk_prescreen_radii(xk_ctx* ctx, const double* coords, double margin, double cell_size,
                       const int32_t* elem_layer, double* rsum_host, int64_t cap,
                       int64_t* n_out, void* stream);

/* Run the pair kernel over all pairs with the given XK_PAIR_* modes.
 * coords_ref (device, n_nodes x 3) plus u (device, nullable) give the
 * deformed coordinates; u_ref (device) is the friction reference.  Covers
 * UpdateGeometryProcess (_manager_process.py:327-403, hysteresis :475-494),
 * HuberContactForceProcess.evaluate/tangent (contact_force/strategy.py:126-254),
 * CoulombReturnMappingProcess.evaluate/tangent (friction/strategy.py:111-164)
 * and the KelvinVoigtCoatingProcess methods (coating/strategy.py:97-379).
 * With the coating on, every geometry update invalidates the tangent pattern
 * (a new compression gives a pair a coating contribution). */
int xk_pairs_run(xk_ctx* ctx, int modes, const double* coords_ref, const double* u,
                 const double* u_ref, double k_pen, double smoothing_delta, double k_t_ratio,
                 double mu_eff, double g_on, double g_off, int allow_deactivation,
                 void* stream);
/* Gather the pair forces staged by the last xk_pairs_run into f (device,
 * float64, 6 n_nodes, overwritten) and/or the staged pair tangent blocks into
 * K_vals (device, pattern order, overwritten, K_cap entries). */
int xk_pairs_gather(xk_ctx* ctx, double* f, void* K_vals, int64_t K_cap, void* stream);

/* (Re)build the tangent pattern: beam blocks plus the blocks of every pair
 * that can contribute within the current increment (status != INACTIVE or
 * p_n > 0).  Called automatically when stale.  Synchronous. */
int xk_pattern_build(xk_ctx* ctx);
int xk_pattern_info(xk_ctx* ctx, int64_t* nnz, int64_t* n_blocks, int64_t* n_contrib_pairs);
/* Scalar CSR of the current pattern (host arrays of ndof+1 and nnz). */
int xk_pattern_export(xk_ctx* ctx, int64_t* indptr, int32_t* indices);

/* Generalized-alpha state for the fused assembly (time_integration/strategy.py:96-232):
 * lumped mass diagonal (device, ndof), u_pred, v_pred, acc_old (device, ndof)
 * are read; acc and vel (device, ndof) are written by xk_assemble_iteration. */
typedef struct xk_dyn {
    const double* M_diag;
    const double* u_pred;
    const double* v_pred;
    const double* acc_old;
    double* acc;
    double* vel;
    double dt;
    double alpha_m;
    double beta;
    double gamma;
} xk_dyn;

/* One fused Newton-attempt assembly: ContactForceAssemblyProcess
 * (contact/solver/_newton_steps.py:79-135) + GeneralizedAlpha correct /
 * effective_residual (_newton_dynamic.py:189-192) + TangentAssemblyProcess
 * (_newton_steps.py:289-313) + effective_stiffness (_newton_dynamic.py:314-315):
 *   narrowphase (frozen active set) -> Huber -> friction -> element f and K ->
 *   R = f_int - (f_c + f_fric) - f_ext, R[fixed] = 0, R += M a_mid,
 *   R[fixed] = 0 -> K = K_beam + K_c - K_fric + (1 - alpha_m) c0 M.
 * fixed_mask (device uint8, ndof); dyn may be NULL (static); dyn->acc and
 * dyn->vel (the per-attempt GeneralizedAlphaProcess.correct state) may both be
 * NULL for a caller that corrects at the increment's end (process.py:549-551),
 * which overwrites them before anything reads them.  R (device),
 * K_vals (device, nullable, K_cap entries), res_norm2 (device double, ||R||^2, nullable),
 * n_active (device int32, #(p_n > 0), nullable). */
int xk_assemble_iteration(xk_ctx* ctx, const double* u, const double* u_ref,
                          const double* f_ext, const uint8_t* fixed_mask, const xk_dyn* dyn,
                          double k_pen, double smoothing_delta, double k_t_ratio, double mu_eff,
                          double* R, void* K_vals, int64_t K_cap, double* res_norm2, int32_t* n_active,
                          void* stream);

/* Linear solve of the attempt on the device, replacing LinearSolveProcess
 * (contact/solver/_newton_steps.py:338-366, SciPy spsolve of
 * (D K D + I - D) du = -D R, D = diag(free dofs)).  Direct element path only
 * (element chains).  Call after xk_assemble_iteration with K_vals on the same
 * ctx, with the same u, fixed_mask and dyn.  The tangent is taken as
 * K = A_b + sum_p U_p G_p U_p^T: the beam part A_b = K_beam + (1 - alpha_m) c0 M
 * (the element kernel re-run, factored along each chain) and the pairs'
 * staged 3x3 tangent blocks G_p with U_p = c_p (x) I3 (Woodbury, a dense
 * capacitance matrix per contact component, LU-factored without row
 * interchanges -- blocked, DGEMM trailing updates -- or, when that meets a
 * zero pivot or leaves an eager refinement above tol, by rocSOLVER's
 * partially pivoted getrf for the rest of the pattern; see xk_solve_set_lu); in the FP64 form
 * this equals the assembled K up to rounding, in the mixed form K holds the
 * same sum rounded to FP32.  The result is refined against that operator:
 * with tol <= 0 exactly max_refine steps, stream-ordered; with tol > 0 until
 * ||D R + A~ du|| <= tol ||D R||, a step no longer halves that residual (its
 * rounding floor, about 1e-12 on strands) or max_refine steps, reading the
 * norm on the host after each step (2 steps reach 1e-11 on strands).  du (device, ndof) is written; res_rel (device
 * double, nullable) receives the final relative residual, n_refine (host,
 * nullable) the steps taken.  A zero pivot sets XK_FLAG_SINGULAR in
 * xk_status_flags.  The first call after a pattern change sets up the solve
 * synchronously. */
int xk_linear_solve(xk_ctx* ctx, const double* u, const uint8_t* fixed_mask, const xk_dyn* dyn,
                    const double* R, double* du, int max_refine, double tol, double* res_rel,
                    int32_t* n_refine, void* stream);
/* Capacitance LU of xk_linear_solve.  kind 0 (default): without row
 * interchanges, falling back to partial pivoting as described there; 1:
 * partial pivoting (rocSOLVER getrf) always; -1: leave unchanged (query).
 * *pivoting (nullable) = 1 if the next solve on the current pattern uses
 * partial pivoting; *fallbacks (nullable) = solves that fell back so far.
 * No counterpart in the reference (SciPy's SuperLU chooses its own pivots). */
int xk_solve_set_lu(xk_ctx* ctx, int kind, int32_t* pivoting, int64_t* fallbacks);
/* Line search, update and convergence norms of the attempt on the device,
 * replacing LineSearchUpdateProcess (contact/solver/_newton_steps.py:396-434)
 * with _ncp_line_search (contact/solver/_utils.py:31-61) and the
 * displacement / energy norms of ConvergenceCheckProcess (:215-243).
 * alpha = 1, 1/2, ... (line_search_steps candidates, 0 = no search): the
 * first alpha whose frozen-contact static residual
 * f_int(u + alpha du) - f_contact - f_ext (fixed dofs zeroed) is finite and
 * below diverge_factor * sqrt(*res_norm2) is taken, else 2^-line_search_steps;
 * with du_norm_cap > 0, alpha du is scaled to at most du_norm_cap * max(||u||, 1);
 * then du = alpha du, u += du.  f_contact is xk_pairs_gather's pair force at
 * the attempt (the reference's -f_c).  status (device, >= 24 doubles):
 * [0] alpha, [1] ||du||, [2] ||u||, [3] |du . R|, [4] index of the accepted
 * candidate (line_search_steps if none), [8 + k] candidate k's residual norm.
 * Direct element path only; stream-ordered, no host synchronisation, no
 * memset (capturable in a HIP graph). */
int xk_newton_update(xk_ctx* ctx, double* u, double* du, const double* R, const double* f_contact,
                     const double* f_ext, const uint8_t* fixed_mask, const double* res_norm2,
                     int line_search_steps, double diverge_factor, double du_norm_cap, double* status,
                     void* stream);

/* ---- Device Newton control: the tests of NewtonDynamicProcess.process
 * (contact/solver/_newton_dynamic.py:160-389) and ConvergenceCheckProcess
 * (contact/solver/_newton_steps.py:182-243) evaluated on the device, so that
 * the attempts of a call can be enqueued back to back with no host read in
 * between.  The control record `ctl` (device doubles, XK_NC_HIST +
 * XK_NC_HSTRIDE * hist_cap) holds the call's state and history.  While a
 * record is attached to the context (xk_newton_control_attach), every kernel
 * of an attempt -- fused assembly, pair force, linear solve, line search and
 * update -- first reads the state word and does nothing once the call has
 * ended (or while an attempt waits on the host), so attempts enqueued past
 * the end are no-ops.  rocBLAS's trailing updates of the capacitance LU are
 * the one exception: they run on stale scratch in a no-op attempt.  The host
 * detaches the record for an eager re-solve (XK_NC_ASSIST), and rebuilds the
 * tangent pattern when a coated attempt's probe found a pair without a slot
 * (XK_NC_REBUILD, xk_coating_probe_control). */
enum {
    XK_NC_STATE = 0,     /* XK_NC_RUNNING .. XK_NC_REBUILD */
    XK_NC_ATT = 1,       /* attempts whose force test ran */
    XK_NC_REASON = 2,    /* XK_NCR_* bits of a stopped call */
    XK_NC_NACT = 3,      /* n_active of the last attempt */
    XK_NC_RATIO = 4,     /* ||R|| / f_ref of the last attempt */
    XK_NC_DU = 5, XK_NC_U = 6, XK_NC_ENERGY = 7,   /* ||du||, ||u||, |du . R| of the last update */
    XK_NC_REL = 8,       /* relative residual of the last linear solve (xk_linear_solve's res_rel) */
    XK_NC_ALPHA = 9,     /* step length of the last update */
    XK_NC_CONSEC = 10, XK_NC_PREV = 11, XK_NC_FREF = 12, XK_NC_EREF = 13, XK_NC_SOLVE_OK = 14,
    XK_NC_RGATE = 15,    /* device-decided refinement: 0 while steps remain */
    XK_NC_TOL_F = 16, XK_NC_TOL_D = 17, XK_NC_FEXT = 18, XK_NC_DYNREF = 19, XK_NC_WIN = 20,
    XK_NC_MAXATT = 21, XK_NC_STOL = 22, XK_NC_HCAP = 23,
    XK_NC_RPREV = 24,    /* last refinement residual that did not stop it */
    XK_NC_RSTEPS = 25,   /* refinement steps of the last solve */
    XK_NC_HIST = 32,     /* attempt k: [ratio, n_active, solve residual, alpha] at XK_NC_HIST + 4 k
                            (the last two NaN when the attempt ended before its solve / update) */
    XK_NC_HSTRIDE = 4
};
enum { XK_NC_RUNNING = 0, XK_NC_CONVERGED = 1, XK_NC_DIVERGED = 2, XK_NC_STOPPED = 3,
       XK_NC_ASSIST = 4,    /* the attempt's solve missed solve_tol: the host re-solves it eagerly */
       XK_NC_REBUILD = 5 }; /* the attempt's coating probe met a pair without a pattern slot: the
                               host rebuilds the pattern and enqueues the attempt again */
enum { XK_NCR_NONFINITE = 1, XK_NCR_DEGENERATE = 2, XK_NCR_SINGULAR = 4, XK_NCR_SOLVE = 8, XK_NCR_LIMIT = 16 };
/* Reset `ctl` for a new call (stream-ordered) and clear the device status
 * word.  f_ref of the force test: with dynamic_ref the call's first residual
 * norm (if > 1e-30), else f_ext_ref_norm (_newton_steps.py:194-213);
 * divergence: divergence_window consecutive ratio increases by more than 1 %
 * unless the ratio is below 0.01, or a ratio above 100 from the 11th attempt
 * on (_newton_dynamic.py:258-285); displacement ||du|| / ||u|| < tol_disp and
 * energy |du . R| / (that of the first update) < 1e-10 after the update
 * (_newton_steps.py:215-243); at most max_attempts attempts; an attempt whose
 * solve leaves a relative residual above solve_tol goes to XK_NC_ASSIST. */
int xk_newton_control_begin(xk_ctx* ctx, double* ctl, int hist_cap, double tol_force, double tol_disp,
                            double f_ext_ref_norm, int dynamic_ref, int divergence_window, int max_attempts,
                            double solve_tol, void* stream);
/* Attach `ctl` to the context (null detaches).  While attached,
 * xk_newton_update also applies the solve check before the line search
 * (XK_FLAG_SINGULAR: stopped; residual above solve_tol: XK_NC_ASSIST) and the
 * displacement, energy and attempt-limit tests after the update, and
 * xk_linear_solve with tol > 0 decides on the device when its refinement
 * stops (same rule as its host loop; the steps after the stop are no-ops, no
 * host read, capturable; the LU without interchanges is not re-checked:
 * a solve it leaves above tol goes to XK_NC_ASSIST). */
int xk_newton_control_attach(xk_ctx* ctx, double* ctl);
/* Force and divergence tests of the attempt after its assembly, on the
 * attached record: res_norm2 and n_active are xk_assemble_iteration's; the
 * device status word stops the call on XK_FLAG_DEGENERATE (checked first, as
 * the reference raises there) or XK_FLAG_NONFINITE (after the tests, as a
 * failed linear solve, _newton_steps.py:360-366). */
int xk_newton_control_force(xk_ctx* ctx, const double* res_norm2, const int32_t* n_active, void* stream);
/* Host decision on the attached record after an eager re-solve or a pattern
 * rebuild: sets the state and ORs `reason` into XK_NC_REASON; XK_NC_RUNNING
 * after XK_NC_ASSIST accepts the attempt's solve (the next update skips the
 * solve check once). */
int xk_newton_control_set(xk_ctx* ctx, int state, int reason, void* stream);

/* Copy of one of the linear solve's internal device buffers (index `which`,
 * layouts in csrc/solve.hip) for tests; *size_bytes receives its size.
 * Synchronous.  No counterpart in the reference. */
int xk_solve_buffer(xk_ctx* ctx, int which, void* host_dst, int64_t nbytes, int64_t* size_bytes);

/* ---- Learned candidate prescreen.  No counterpart in the reference: an
 * 8-64-64-1 BF16 MLP (inference on the matrix cores) chooses which candidate
 * pairs the fused assembly's pair kernel visits in the Newton iterations of an
 * increment; every contributing pair (status != INACTIVE or p_n > 0) is always
 * visited, so R, K_T and n_active equal the full path's.  Parameters: FP32,
 * W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3 (xk_mlp_n_params()). */
int xk_lp_set_params(xk_ctx* ctx, const float* params_host, int n_params, float threshold);
/* Features of every pair at x = coords_ref + u (device), elem_layer (device
 * int32 per element, nullable) scaled by layer_scale; features_out (device
 * float, n_pairs x 8, nullable) receives a copy for logging / training.  Call
 * after the increment's geometry update; remembers each pair's gap for the
 * next increment's gap-change feature. */
int xk_lp_features(xk_ctx* ctx, const double* coords_ref, const double* u, const int32_t* elem_layer,
                   float layer_scale, float* features_out, void* stream);
/* Inference + watch / prune lists of the current pattern; switches the fused
 * assembly to the watch list until the next xk_lp_verify / xk_lp_off, pattern
 * rebuild or pair-list change.  Synchronous. */
int xk_lp_select(xk_ctx* ctx, int64_t* n_watch, void* stream);
/* Verification pass: the pruned pairs get the narrowphase and frame transport
 * of every fused assembly since xk_lp_select, in order (each assembly over the
 * watch list records its translational displacement, up to 64 of them; the
 * frame transport depends on the path, so this leaves every pair's state
 * bitwise as the full path's assemblies would), active set frozen;
 * n_miss = #(pruned pairs with gap <= g_on at the last geometry), plus 1 if
 * more than 64 assemblies ran (the caller then re-runs the call on the full
 * path).  coords_ref is the reference coordinates (device, n_nodes x 3); u
 * is not read.  Switches the watch list off.  Synchronous. */
int xk_lp_verify(xk_ctx* ctx, const double* coords_ref, const double* u, double g_on, int64_t* n_miss,
                void* stream);
int xk_lp_off(xk_ctx* ctx);
/* Copy of the whole pair state on the device, and its restore (the pair list
 * must not have changed in between); used to re-run a Newton call on the full
 * path after a prescreen miss.  Synchronous. */
int xk_pairs_snapshot(xk_ctx* ctx);
int xk_pairs_restore(xk_ctx* ctx);
/* Stand-alone network kernels (device pointers; no context).  Forward: logits
 * of n feature rows (FP32, rounded to BF16 on load).  Gradient: mean weighted
 * logistic loss w+ y softplus(-z) + (1 - y) softplus(z) and its gradient with
 * respect to the FP32 parameters (grad_and_loss: n_params gradient entries,
 * then the loss), deterministic; work holds xk_mlp_grad_work(n) floats.  Adam:
 * one bias-corrected step (step counts from 1). */
int xk_mlp_n_params(void);
int64_t xk_mlp_grad_work(int64_t n);
int xk_mlp_forward(const float* X, int64_t n, const float* params, float* logits, void* stream);
int xk_mlp_grad(const float* X, const float* y, int64_t n, const float* params, float pos_weight,
                float* grad_and_loss, float* work, int64_t work_len, void* stream);
int xk_mlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr, float beta1,
                float beta2, float eps, void* stream);
/* xk_mlp_grad without the workspace argument and without any allocation: the
 * per-workgroup gradient partials go to a fixed device array of the library
 * (256 x 4802 floats, 4.9 MB, held by every GPU the library runs on until the
 * process ends).  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP-graph capture, training loops that forbid
 * hipMalloc).  Any n >= 1 fits (the gradient kernel's grid stops at 256
 * workgroups); a layout that needed more would be rejected with an error.
 * Calls no hipMalloc / hipFree; the first call on a device resolves the
 * array's address there (make it outside graph capture).  Calls that overlap
 * on one GPU (different streams, or threads) share the array and race;
 * calls ordered on one stream do not.  Results are bitwise xk_mlp_grad's. */
int xk_mlp_grad_noalloc(const float* X, const float* y, int64_t n, const float* params, float pos_weight,
                        float* grad_and_loss, void* stream);

/* Per-element friction coefficients (host, n_elems, >= 0; NULL = off): a pair
 * uses mu_eff x mu_elem[elem_a] instead of mu_eff.  No counterpart in the
 * reference, whose solve has one coefficient: it lets independent solves with
 * their own coefficients (mu_eff = 1) share one context and one launch. */
int xk_ctx_set_friction(xk_ctx* ctx, const double* mu_elem);

/* Kelvin-Voigt coating pair terms.  Replaces _create_coating_strategy and
 * KelvinVoigtCoatingProcess (xkep_cae/contact/coating/strategy.py:80-400) with
 * the coating branches of UpdateGeometryProcess and the active-set update
 * (xkep_cae/contact/_manager_process.py:351-360, :486-494): on when
 * stiffness > 0 (the config's coating_stiffness / coating_damping /
 * coating_mu / coating_k_t_ratio, xkep_cae/contact/_contact_pair.py:164-167).
 * core_radii: per element, host (NULL: the contact radii, which is what the
 * reference's solver passes, contact/solver/process.py:229); pairs created
 * later and the existing ones take them.  Every pair-kernel call then uses
 * the core gap, writes the compression, and the fused assembly adds the
 * coating force, coating friction and their stiffness.  Synchronous. */
int xk_ctx_set_coating(xk_ctx* ctx, double stiffness, double damping, double mu, double k_t_ratio,
                       const double* core_radii);
/* The dashpot's time step of the following assemblies,
 * max(load_frac - load_frac_prev, 1e-15) (contact/solver/_newton_steps.py:112, :297). */
int xk_coating_set_dt(xk_ctx* ctx, double coat_dt);
/* Compression of every pair at x = X + u, nothing else written: pairs with a
 * compression > 0 are marked for the next pattern build; n_unplaced = how
 * many of them hold no slot of the current pattern (then the pattern is
 * rebuilt by the next assembly call).  A frozen-INACTIVE pair can start
 * pressing its coating within an increment, and the reference applies the
 * coating force to it regardless of status; call this before an assembly
 * whenever the coating is on.  No counterpart in the reference (whose
 * assembly has no fixed pattern).  Synchronous. */
int xk_coating_probe(xk_ctx* ctx, const double* u, int64_t* n_unplaced, void* stream);
/* xk_coating_probe on the attached Newton control record, stream-ordered with
 * no host read (capturable): a compressed pair without a slot sets the
 * record's state to XK_NC_REBUILD, so the rest of the attempt is a no-op; the
 * host then calls xk_coating_probe (which marks the pattern stale), rebuilds
 * the pattern, sets XK_NC_RUNNING and enqueues the attempt again.  Nothing to
 * do while the call is not running.  No counterpart in the reference. */
int xk_coating_probe_control(xk_ctx* ctx, const double* u, void* stream);
/* End of a converged step: coating_compression_prev <- coating_compression
 * (contact/solver/process.py:539-547).  Stream-ordered. */
int xk_coating_commit(xk_ctx* ctx, void* stream);
/* Coating state of the pair list (_contact_pair.py:57-62, core radii :79-80),
 * host arrays of n_pairs (z_t: n_pairs x 2), each nullable.  Synchronous. */
int xk_pairs_get_coating(xk_ctx* ctx, double* core_a, double* core_b, double* compression,
                         double* compression_prev, double* z_t, int8_t* stick, double* q_trial_norm,
                         double* dissipation);
int xk_pairs_set_coating(xk_ctx* ctx, const double* compression, const double* compression_prev,
                         const double* z_t, const int8_t* stick, const double* q_trial_norm,
                         const double* dissipation);

/* Launch shapes and the broadphase grid (no counterpart in the reference;
 * set by the autotuner): grid cell = cell_factor x mean box extent when a call
 * passes cell_size <= 0 (the candidate set does not depend on it), pair-kernel
 * workgroup 64/128/256/512 lanes, gather workgroup 128/256/512 lanes.  A HIP
 * graph captured before a change keeps the old shapes. */
typedef struct xk_tuning {
    double cell_factor;
    int pair_block;
    int gather_block;
} xk_tuning;
int xk_ctx_set_tuning(xk_ctx* ctx, const xk_tuning* tuning);
int xk_ctx_get_tuning(const xk_ctx* ctx, xk_tuning* tuning);

/* Device status word of the last calls (XK_FLAG_* bits); synchronous. */
int xk_status(xk_ctx* ctx, int* flags);

/* Per-stage timing of xk_assemble_iteration with HIP events recorded on the
 * call's stream: begin allocates events for up to max_calls calls; end
 * synchronises and returns in stage_ms4 the summed milliseconds of the four
 * stages [pair kernel, element kernel, gather kernel (residual + tangent),
 * finish kernel (norm + active count)] and the number of calls recorded.  Profiled calls run the three stages in sequence
 * (unprofiled calls overlap the element and pair kernels on two streams).
 * Not for use under graph capture. */
int xk_profile_begin(xk_ctx* ctx, int max_calls);
int xk_profile_end(xk_ctx* ctx, double* stage_ms4, int* n_calls);

#ifdef __cplusplus
}
#endif
#endif /* XKAMD_H */
