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
// Reductions run in one workgroup in a fixed order (deterministic, no atomics,
// no memset).

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

__global__ void nt_axpy_kernel(int n, const double* __restrict__ u, const double* __restrict__ du, int k,
                               double* __restrict__ out) {
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

// ||f_int - f_contact - f_ext|| over free dofs -> status[NS_NORM0 + k] (inf if non-finite)
__global__ __launch_bounds__(NR) void nt_trial_norm_kernel(int n, const double* __restrict__ f_int,
                                                           const double* __restrict__ f_contact,
                                                           const double* __restrict__ f_ext,
                                                           const uint8_t* __restrict__ fixed, int k,
                                                           double* __restrict__ status) {
    __shared__ double red[NR];
    double s = 0.0;
    int bad = 0;
    for (int i = threadIdx.x; i < n; i += NR) {
        double r = (f_int[i] - f_contact[i]) - f_ext[i];
        if (fixed[i]) r = 0.0;
        if (!isfinite(r)) bad = 1;
        s = fma(r, r, s);
    }
    const double tot = nt_block_sum(s, red);
    const double nb = nt_block_sum((double)bad, red);
    if (threadIdx.x == 0) status[NS_NORM0 + k] = nb > 0.0 ? HUGE_VAL : sqrt(tot);
}

// alpha from the candidates' norms, then the cap on ||alpha du||
__global__ __launch_bounds__(NR) void nt_select_kernel(int n, int max_steps, double diverge_factor, double cap,
                                                       const double* __restrict__ res_norm2,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ du,
                                                       double* __restrict__ status) {
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
__global__ void nt_update_kernel(int n, double* __restrict__ u, double* __restrict__ du,
                                 const double* __restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = status[NS_ALPHA] * du[i];
    du[i] = d;
    u[i] += d;
}

__global__ __launch_bounds__(NR) void nt_norms_kernel(int n, const double* __restrict__ u,
                                                      const double* __restrict__ du, const double* __restrict__ R,
                                                      double* __restrict__ status) {
    __shared__ double red[NR];
    double sd = 0.0, su = 0.0, se = 0.0;
    for (int i = threadIdx.x; i < n; i += NR) {
        sd = fma(du[i], du[i], sd);
        su = fma(u[i], u[i], su);
        se = fma(du[i], R[i], se);
    }
    const double a = nt_block_sum(sd, red), b = nt_block_sum(su, red), c = nt_block_sum(se, red);
    if (threadIdx.x == 0) {
        status[NS_DU] = sqrt(a);
        status[NS_U] = sqrt(b);
        status[NS_ENERGY] = fabs(c);
    }
}

static unsigned nt_grid(int n) { return (unsigned)std::max(1, (n + 255) / 256); }

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
    if (nt_try.ensure(sizeof(double) * (size_t)n) != hipSuccess) {
        set_error("xk_newton_update: out of device memory");
        return -1;
    }
    double* ut = (double*)nt_try.p;
    for (int k = 0; k < max_steps; ++k) {
        hipLaunchKernelGGL(nt_axpy_kernel, dim3(nt_grid(n)), dim3(256), 0, st, n, (const double*)u,
                           (const double*)du, k, ut);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = launch_elements(ut, true, nullptr, 0.0, nullptr, st);
        if (e != hipSuccess) {
            set_error(std::string("xk_newton_update: ") + hipGetErrorString(e));
            return -1;
        }
        hipLaunchKernelGGL(nt_trial_norm_kernel, dim3(1), dim3(NR), 0, st, n, (const double*)d_fint.p, f_contact,
                           f_ext, fixed, k, status);
    }
    hipLaunchKernelGGL(nt_select_kernel, dim3(1), dim3(NR), 0, st, n, max_steps, diverge_factor, cap, res_norm2,
                       (const double*)u, (const double*)du, status);
    hipLaunchKernelGGL(nt_update_kernel, dim3(nt_grid(n)), dim3(256), 0, st, n, u, du, (const double*)status);
    hipLaunchKernelGGL(nt_norms_kernel, dim3(1), dim3(NR), 0, st, n, (const double*)u, (const double*)du, R, status);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("xk_newton_update: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

}  // namespace xk
