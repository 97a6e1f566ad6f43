// Line search, update and convergence norms of one Newton attempt on the
// device, stream-ordered and free of host synchronisation (so that the
// attempt can be captured in a graph):
//
//   LineSearchUpdateProcess (contact/solver/_newton_steps.py:396-434) with
//   _ncp_line_search (contact/solver/_utils.py:31-61): alpha = 1, 1/2, ...;
//   accept the first alpha whose frozen-contact static residual
//   f_int(u + alpha du) - f_contact - f_ext (fixed dofs zeroed) is finite and
//   below diverge_factor * ||R||; after max_steps rejections alpha = 2^-max_steps;
//   the optional cap on ||alpha du|| relative to max(||u||, 1); u += alpha du;
//   then ||du||, ||u|| and |du . R| for ConvergenceCheckProcess
//   (_newton_steps.py:215-243).
//
// Every candidate alpha is evaluated (one element force pass and one residual
// norm each) and the first accepted one is selected on the device: the result
// is the sequential search's, without a host round trip per candidate.
// The candidates' residual norms and the update's norms are reduced in two
// launches over the chip, the optional step cap's norms in one workgroup, each
// in a fixed order (deterministic, no atomics, no memset).
//
// The device Newton control (the nc_* kernels) evaluates the driver's tests
// into a per-call control record: force convergence and divergence after the
// assembly, the solve check before the line search, displacement and energy
// convergence and the attempt limit after the update.  Each is a single-lane
// kernel with the host driver's arithmetic, so the device takes the branches
// the host would; every attempt kernel of the library reads the record's state
// word first and does nothing once the call has ended.

#include "../../include/xkamd.h"
#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"

#include <algorithm>

namespace xk {

constexpr int NR = 1024;   // reduction workgroup
constexpr int NT_MAX_LS = 16;

// uniform layout of the status buffer (doubles)
enum { NS_ALPHA = 0, NS_DU = 1, NS_U = 2, NS_ENERGY = 3, NS_LS = 4, NS_NORM0 = 8 };

__global__ void nt_axpy_kernel(const double* __restrict__ stop, int n, const double* __restrict__ u,
                               const double* __restrict__ du, int k, double* __restrict__ out) {
    if (!nc_go(stop)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = u[i] + ldexp(1.0, -k) * du[i];
}

__device__ __forceinline__ double nt_block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = NR / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// Norms over the ndof vectors in two launches: NP_BLOCK lanes per workgroup,
// each workgroup's partial sums (in lane order, then a fixed tree) to part[],
// then one workgroup folding the partials in index order -- deterministic,
// no atomics, and spread over the chip instead of one workgroup.
constexpr int NP_BLOCK = 256;
constexpr int NP_PER_LANE = 4;
static int np_blocks(int n) { return std::max(1, (n + NP_BLOCK * NP_PER_LANE - 1) / (NP_BLOCK * NP_PER_LANE)); }

// sums of v[0..NV) over the workgroup (lane order, then a fixed tree), valid in lane 0
template <int NV>
__device__ __forceinline__ void np_block_sum(double (&v)[NV]) {
    __shared__ double red[NV][NP_BLOCK];
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = NP_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < NV; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = red[q][0];
}

// partials part[q * gridDim + block] of sum r^2 (q = 0) and #(non-finite r)
// (q = 1), r = f_int - f_contact - f_ext on free dofs
__global__ __launch_bounds__(NP_BLOCK) void nt_trial_part_kernel(const double* __restrict__ stop, int n,
                                                                 const double* __restrict__ f_int,
                                                                 const double* __restrict__ f_contact,
                                                                 const double* __restrict__ f_ext,
                                                                 const uint8_t* __restrict__ fixed,
                                                                 double* __restrict__ part) {
    if (!nc_go(stop)) return;
    double v[2] = {0.0, 0.0};
    const int i0 = blockIdx.x * NP_BLOCK * NP_PER_LANE + threadIdx.x;
#pragma unroll
    for (int q = 0; q < NP_PER_LANE; ++q) {
        const int i = i0 + q * NP_BLOCK;
        if (i < n) {
            double r = (f_int[i] - f_contact[i]) - f_ext[i];
            if (fixed[i]) r = 0.0;
            if (!isfinite(r)) v[1] += 1.0;
            v[0] = fma(r, r, v[0]);
        }
    }
    np_block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = v[0];
        part[gridDim.x + blockIdx.x] = v[1];
    }
}

// ||r|| of candidate k from the nb partials -> status[NS_NORM0 + k] (inf if
// any entry is non-finite)
__global__ __launch_bounds__(NP_BLOCK) void nt_trial_norm_kernel(const double* __restrict__ stop, int nb,
                                                                 const double* __restrict__ part, int k,
                                                                 double* __restrict__ status) {
    if (!nc_go(stop)) return;
    double v[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += NP_BLOCK) {
        v[0] += part[b];
        v[1] += part[nb + b];
    }
    np_block_sum<2>(v);
    if (threadIdx.x == 0) status[NS_NORM0 + k] = v[1] > 0.0 ? HUGE_VAL : sqrt(v[0]);
}

// alpha from the candidates' norms, then the cap on ||alpha du||
__global__ __launch_bounds__(NR) void nt_select_kernel(const double* __restrict__ stop, int n, int max_steps,
                                                       double diverge_factor, double cap,
                                                       const double* __restrict__ res_norm2,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ du,
                                                       double* __restrict__ status) {
    if (!nc_go(stop)) return;
    __shared__ double red[NR];
    double alpha = 1.0;
    int used = max_steps;
    if (max_steps > 0) {
        const double thr = diverge_factor * fmax(sqrt(*res_norm2), 1e-30);
        alpha = ldexp(1.0, -max_steps);
        for (int k = 0; k < max_steps; ++k)
            if (status[NS_NORM0 + k] < thr) {   // finite and below the threshold
                alpha = ldexp(1.0, -k);
                used = k;
                break;
            }
    }
    if (cap > 0.0) {
        double sd = 0.0, su = 0.0;
        for (int i = threadIdx.x; i < n; i += NR) {
            const double a = alpha * du[i];
            sd = fma(a, a, sd);
            su = fma(u[i], u[i], su);
        }
        const double dn = sqrt(nt_block_sum(sd, red));
        const double un = fmax(sqrt(nt_block_sum(su, red)), 1.0);
        if (dn > cap * un) alpha *= cap * un / dn;
    }
    if (threadIdx.x == 0) {
        status[NS_ALPHA] = alpha;
        status[NS_LS] = (double)used;
    }
}

// du *= alpha; u += du
__global__ void nt_update_kernel(const double* __restrict__ stop, int n, double* __restrict__ u,
                                 double* __restrict__ du, const double* __restrict__ status) {
    if (!nc_go(stop)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = status[NS_ALPHA] * du[i];
    du[i] = d;
    u[i] += d;
}

// partials of ||du||^2, ||u||^2 and du . R
__global__ __launch_bounds__(NP_BLOCK) void nt_norms_part_kernel(const double* __restrict__ stop, int n,
                                                                 const double* __restrict__ u,
                                                                 const double* __restrict__ du,
                                                                 const double* __restrict__ R,
                                                                 double* __restrict__ part) {
    if (!nc_go(stop)) return;
    double v[3] = {0.0, 0.0, 0.0};
    const int i0 = blockIdx.x * NP_BLOCK * NP_PER_LANE + threadIdx.x;
#pragma unroll
    for (int q = 0; q < NP_PER_LANE; ++q) {
        const int i = i0 + q * NP_BLOCK;
        if (i < n) {
            v[0] = fma(du[i], du[i], v[0]);
            v[1] = fma(u[i], u[i], v[1]);
            v[2] = fma(du[i], R[i], v[2]);
        }
    }
    np_block_sum<3>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 3; ++q) part[q * gridDim.x + blockIdx.x] = v[q];
}

// ||du||, ||u||, |du . R| from the nb partials
__global__ __launch_bounds__(NP_BLOCK) void nt_norms_kernel(const double* __restrict__ stop, int nb,
                                                            const double* __restrict__ part,
                                                            double* __restrict__ status) {
    if (!nc_go(stop)) return;
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += NP_BLOCK)
#pragma unroll
        for (int q = 0; q < 3; ++q) v[q] += part[q * nb + b];
    np_block_sum<3>(v);
    if (threadIdx.x == 0) {
        status[NS_DU] = sqrt(v[0]);
        status[NS_U] = sqrt(v[1]);
        status[NS_ENERGY] = fabs(v[2]);
    }
}

static unsigned nt_grid(int n) { return (unsigned)std::max(1, (n + 255) / 256); }

// ------------------------------------------------------------ device control
//
// The control record of a call (XK_NC_* layout of include/xkamd.h) is read
// and written by single-lane kernels in stream order, so every test sees the
// values of the attempt before it; the arithmetic (sqrt, division, the 1.01
// and 0.01 / 100 thresholds, NaN comparisons) is the host driver's, so the
// device takes the same branch for the same numbers.

static_assert(XK_NC_RUNNING == 0, "nc_go (xk_common.h) tests this value");

__global__ void nc_begin_kernel(double* __restrict__ c, int hcap, double tol_f, double tol_d, double fext,
                                int dynref, int win, int maxatt, double stol, int* __restrict__ flags) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < XK_NC_HIST; ++k) c[k] = 0.0;
    c[XK_NC_STATE] = XK_NC_RUNNING;
    c[XK_NC_PREV] = HUGE_VAL;      // prev_ratio = inf
    c[XK_NC_EREF] = -1.0;          // energy_ref not set
    c[XK_NC_TOL_F] = tol_f;
    c[XK_NC_TOL_D] = tol_d;
    c[XK_NC_FEXT] = fext;
    c[XK_NC_DYNREF] = dynref;
    c[XK_NC_WIN] = win;
    c[XK_NC_MAXATT] = maxatt;
    c[XK_NC_STOL] = stol;
    c[XK_NC_HCAP] = hcap;
    *flags = 0;
}

__device__ __forceinline__ void nc_stop(double* c, int state, int reason) {
    c[XK_NC_STATE] = state;
    c[XK_NC_REASON] = (double)((int)c[XK_NC_REASON] | reason);
}

// ConvergenceCheckProcess, force part (_newton_steps.py:194-213), and the
// divergence tests (_newton_dynamic.py:258-285), after the attempt's assembly
__global__ void nc_force_kernel(double* __restrict__ c, const double* __restrict__ res_norm2,
                                const int* __restrict__ n_active, const int* __restrict__ flags) {
    if (threadIdx.x != 0 || c[XK_NC_STATE] != XK_NC_RUNNING) return;
    const int att = (int)c[XK_NC_ATT];
    const int fl = *flags;
    const double res = sqrt(*res_norm2);
    const bool dyn = c[XK_NC_DYNREF] != 0.0;
    const double incr = c[XK_NC_FREF];
    double f_ref = incr > 1e-30 ? incr : c[XK_NC_FEXT];
    if (dyn && att == 0 && res > 1e-30) f_ref = res;
    if (att == 0 && dyn) c[XK_NC_FREF] = f_ref;
    const double ratio = res / f_ref;
    const int na = *n_active;
    c[XK_NC_RATIO] = ratio;
    c[XK_NC_NACT] = na;
    if (att < (int)c[XK_NC_HCAP]) {
        double* h = c + XK_NC_HIST + XK_NC_HSTRIDE * att;
        h[0] = ratio;
        h[1] = na;
        h[2] = NAN;   // solve residual and step length: NaN until the attempt gets that far
        h[3] = NAN;
    }
    c[XK_NC_ATT] = att + 1;
    // a degenerate element: the reference raises in the element routine
    if (fl & XK_FLAG_DEGENERATE) { nc_stop(c, XK_NC_STOPPED, XK_NCR_DEGENERATE); return; }
    if (ratio < c[XK_NC_TOL_F]) { c[XK_NC_STATE] = XK_NC_CONVERGED; return; }
    int consec = (int)c[XK_NC_CONSEC];
    if (att > 0 && ratio > c[XK_NC_PREV] * 1.01) consec += 1;
    else consec = 0;
    c[XK_NC_CONSEC] = consec;
    c[XK_NC_PREV] = ratio;
    if ((consec >= (int)c[XK_NC_WIN] && !(ratio < 0.01)) || (att >= 10 && ratio > 100.0)) {
        c[XK_NC_STATE] = XK_NC_DIVERGED;
        return;
    }
    // a non-finite system has no solution: LinearSolveProcess success=False (_newton_steps.py:360-366)
    if ((fl & XK_FLAG_NONFINITE) || !isfinite(res)) nc_stop(c, XK_NC_STOPPED, XK_NCR_NONFINITE);
}

// after the linear solve, before the line search: a zero pivot stops the call;
// a residual above the tolerance hands the attempt to the host (XK_NC_ASSIST),
// unless the host has just accepted this attempt's solve
__global__ void nc_solve_kernel(double* __restrict__ c, const int* __restrict__ flags) {
    if (threadIdx.x != 0 || c[XK_NC_STATE] != XK_NC_RUNNING) return;
    const int att = (int)c[XK_NC_ATT] - 1;
    if (att >= 0 && att < (int)c[XK_NC_HCAP]) c[XK_NC_HIST + XK_NC_HSTRIDE * att + 2] = c[XK_NC_REL];
    if (c[XK_NC_SOLVE_OK] != 0.0) {
        c[XK_NC_SOLVE_OK] = 0.0;
        return;
    }
    if (*flags & XK_FLAG_SINGULAR) { nc_stop(c, XK_NC_STOPPED, XK_NCR_SINGULAR); return; }
    if (!(c[XK_NC_REL] <= c[XK_NC_STOL])) c[XK_NC_STATE] = XK_NC_ASSIST;   // NaN included
}

// ConvergenceCheckProcess, displacement and energy part (_newton_steps.py:215-243),
// after the update, and the attempt limit
__global__ void nc_after_kernel(double* __restrict__ c, const double* __restrict__ status) {
    if (threadIdx.x != 0 || c[XK_NC_STATE] != XK_NC_RUNNING) return;
    const double dn = status[NS_DU], un = status[NS_U], en = status[NS_ENERGY];
    c[XK_NC_DU] = dn;
    c[XK_NC_U] = un;
    c[XK_NC_ENERGY] = en;
    c[XK_NC_ALPHA] = status[NS_ALPHA];
    const int att = (int)c[XK_NC_ATT] - 1;
    if (att >= 0 && att < (int)c[XK_NC_HCAP]) c[XK_NC_HIST + XK_NC_HSTRIDE * att + 3] = status[NS_ALPHA];
    if (un > 1e-30 && dn / un < c[XK_NC_TOL_D]) { c[XK_NC_STATE] = XK_NC_CONVERGED; return; }
    double eref = c[XK_NC_EREF];
    if (eref < 0.0) {
        eref = en > 1e-30 ? en : 1.0;
        c[XK_NC_EREF] = eref;
    }
    if (eref > 1e-30 && en / eref < 1e-10) { c[XK_NC_STATE] = XK_NC_CONVERGED; return; }
    if (att + 1 >= (int)c[XK_NC_MAXATT]) nc_stop(c, XK_NC_STOPPED, XK_NCR_LIMIT);
}

__global__ void nc_set_kernel(double* __restrict__ c, int state, int reason) {
    if (threadIdx.x != 0) return;
    const bool assisted = c[XK_NC_STATE] == XK_NC_ASSIST;
    c[XK_NC_STATE] = state;
    c[XK_NC_REASON] = (double)((int)c[XK_NC_REASON] | reason);
    if (state == XK_NC_RUNNING && assisted) c[XK_NC_SOLVE_OK] = 1.0;   // the host's re-solve is accepted
}

int Ctx::newton_control_begin(double* ctl, int hist_cap, double tol_force, double tol_disp, double f_ext_ref_norm,
                              int dynamic_ref, int divergence_window, int max_attempts, double solve_tol,
                              hipStream_t st) {
    if (hist_cap < 0 || max_attempts < 1) {
        set_error("xk_newton_control_begin: invalid history capacity or attempt limit");
        return -1;
    }
    hipLaunchKernelGGL(nc_begin_kernel, dim3(1), dim3(64), 0, st, ctl, hist_cap, tol_force, tol_disp,
                       f_ext_ref_norm, dynamic_ref, divergence_window, max_attempts, solve_tol, (int*)d_flags.p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("xk_newton_control_begin: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int Ctx::newton_control_force(const double* res_norm2, const int* n_active, hipStream_t st) {
    if (!nstop) {
        set_error("xk_newton_control_force: no control record attached");
        return -1;
    }
    hipLaunchKernelGGL(nc_force_kernel, dim3(1), dim3(64), 0, st, nstop, res_norm2, n_active,
                       (const int*)d_flags.p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("xk_newton_control_force: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int Ctx::newton_control_set(int state, int reason, hipStream_t st) {
    if (!nstop || state < XK_NC_RUNNING || state > XK_NC_ASSIST) {
        set_error("xk_newton_control_set: no control record attached or invalid state");
        return -1;
    }
    hipLaunchKernelGGL(nc_set_kernel, dim3(1), dim3(64), 0, st, nstop, state, reason);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("xk_newton_control_set: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int Ctx::newton_update(double* u, double* du, const double* R, const double* f_contact, const double* f_ext,
                       const uint8_t* fixed, const double* res_norm2, int max_steps, double diverge_factor,
                       double cap, double* status, hipStream_t st) {
    if (!direct) {
        set_error("xk_newton_update: needs element chains (the direct element path)");
        return -1;
    }
    if (max_steps < 0 || max_steps > NT_MAX_LS) {
        set_error("xk_newton_update: line-search steps must lie in [0, 16]");
        return -1;
    }
    const int n = 6 * n_nodes;
    const int nb = np_blocks(n);
    if (nt_try.ensure(sizeof(double) * (size_t)n) != hipSuccess ||
        nt_part.ensure(sizeof(double) * 3 * (size_t)nb) != hipSuccess) {
        set_error("xk_newton_update: out of device memory");
        return -1;
    }
    double* part = (double*)nt_part.p;
    const double* stop = nstop;
    if (nstop) hipLaunchKernelGGL(nc_solve_kernel, dim3(1), dim3(64), 0, st, nstop, (const int*)d_flags.p);
    double* ut = (double*)nt_try.p;
    for (int k = 0; k < max_steps; ++k) {
        hipLaunchKernelGGL(nt_axpy_kernel, dim3(nt_grid(n)), dim3(256), 0, st, stop, n, (const double*)u,
                           (const double*)du, k, ut);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = launch_elements(ut, true, nullptr, 0.0, nullptr, st);
        if (e != hipSuccess) {
            set_error(std::string("xk_newton_update: ") + hipGetErrorString(e));
            return -1;
        }
        hipLaunchKernelGGL(nt_trial_part_kernel, dim3(nb), dim3(NP_BLOCK), 0, st, stop, n, (const double*)d_fint.p,
                           f_contact, f_ext, fixed, part);
        hipLaunchKernelGGL(nt_trial_norm_kernel, dim3(1), dim3(NP_BLOCK), 0, st, stop, nb, (const double*)part, k,
                           status);
    }
    hipLaunchKernelGGL(nt_select_kernel, dim3(1), dim3(NR), 0, st, stop, n, max_steps, diverge_factor, cap,
                       res_norm2, (const double*)u, (const double*)du, status);
    hipLaunchKernelGGL(nt_update_kernel, dim3(nt_grid(n)), dim3(256), 0, st, stop, n, u, du,
                       (const double*)status);
    hipLaunchKernelGGL(nt_norms_part_kernel, dim3(nb), dim3(NP_BLOCK), 0, st, stop, n, (const double*)u,
                       (const double*)du, R, part);
    hipLaunchKernelGGL(nt_norms_kernel, dim3(1), dim3(NP_BLOCK), 0, st, stop, nb, (const double*)part, status);
    if (nstop) hipLaunchKernelGGL(nc_after_kernel, dim3(1), dim3(64), 0, st, nstop, (const double*)status);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("xk_newton_update: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

}  // namespace xk
