// Deterministic gathers of staged element / pair contributions into the
// global residual and the block-CSR tangent, plus the Generalized-alpha
// effective terms and the FP64 residual norm.
//
// Every output entry is summed by one lane over a contribution list fixed at
// pattern-build time (elements in element order, then pairs in pair order), so
// results are bitwise reproducible run to run (no floating-point atomics).
//
// Residual (per node, 6 DOFs) follows ContactForceAssemblyProcess
// (xkep_cae/contact/solver/_newton_steps.py:124-129) and the dynamic steps
// of _newton_dynamic.py:189-192 with GeneralizedAlphaProcess.correct /
// effective_residual (time_integration/strategy.py:178-216):
//   R = f_int - (f_c + f_fric) - f_ext ; R[fixed] = 0
//   acc = c0 (u - u_pred) ; vel = v_pred + dt gamma acc
//   R += M ((1 - alpha_m) acc + alpha_m acc_old) ; R[fixed] = 0
// Tangent (per 3x3 block) follows TangentAssemblyProcess (:289-313) and
// effective_stiffness (time_integration/strategy.py:190-203):
//   K = K_beam + K_c - K_fric + (1 - alpha_m) c0 M.

#include "xk_common.h"
#include "xk_structs.h"
#include <type_traits>

namespace xk {

// mode: 0 = f_int only, 1 = pair force only, 2 = full residual
__global__ void __launch_bounds__(256)
node_gather_kernel(int n_nodes, int n_elems, int n_sinc, NodeLists L, const double* __restrict__ fe,
                   const double* __restrict__ sd, int mode, const double* __restrict__ u,
                   const double* __restrict__ f_ext, const uint8_t* __restrict__ fixed, DynArgs dy,
                   double* __restrict__ out, double* __restrict__ partial, int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    if (i < n_nodes) {
        double fb[6] = {0, 0, 0, 0, 0, 0};
        double fp[3] = {0, 0, 0};
        if (mode != 1) {
            for (int q = L.bptr[i]; q < L.bptr[i + 1]; ++q) {
                int code = L.bsrc[q];
                int e = code >> 1, kn = code & 1;
#pragma unroll
                for (int d = 0; d < 6; ++d) fb[d] += fe[(size_t)(6 * kn + d) * n_elems + e];
            }
        }
        if (mode != 0 && L.pptr) {
            for (int q = L.pptr[i]; q < L.pptr[i + 1]; ++q) {
                int code = L.psrc[q];
                int p = code >> 2, k = code & 3;
                double ck = sd[(size_t)k * n_sinc + p];
#pragma unroll
                for (int d = 0; d < 3; ++d) fp[d] += ck * sd[(size_t)(4 + d) * n_sinc + p];
            }
        }
        const size_t o = 6 * (size_t)i;
        if (mode == 0) {
#pragma unroll
            for (int d = 0; d < 6; ++d) out[o + d] = fb[d];
        } else if (mode == 1) {
#pragma unroll
            for (int d = 0; d < 6; ++d) out[o + d] = d < 3 ? fp[d] : 0.0;
        } else {
            int bad = 0;
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                double r = (fb[d] + (d < 3 ? -fp[d] : -0.0)) - f_ext[o + d];
                if (fixed[o + d]) r = 0.0;
                if (dy.on) {
                    double ac = dy.c0 * (u[o + d] - dy.u_pred[o + d]);
                    dy.acc[o + d] = ac;
                    dy.vel[o + d] = dy.v_pred[o + d] + dy.dt * dy.gamma * ac;
                    double amid = (1.0 - dy.alpha_m) * ac + dy.alpha_m * dy.acc_old[o + d];
                    r = r + dy.M[o + d] * amid;
                    if (fixed[o + d]) r = 0.0;
                }
                out[o + d] = r;
                ss += r * r;
                bad |= !isfinite(r);
            }
            if (bad) atomicOr(flags, FLAG_NONFINITE);
        }
    }
    if (partial) {
        // deterministic block partial: fixed-shape tree in LDS
        __shared__ double red[256];
        red[threadIdx.x] = ss;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
    }
}

__global__ void __launch_bounds__(256)
sum_partials_kernel(int n, const double* __restrict__ partial, double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) s += partial[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// beam_sign / pair_sign: 0 excludes the term; mass_coef multiplies M on the diagonal.
template <class KT, class ST, class GT>
__global__ void __launch_bounds__(256)
block_gather_kernel(int n_blocks, int n_elems, int n_sinc, BlockLists B, const ST* __restrict__ Ke,
                    const double* __restrict__ sd, const GT* __restrict__ sg, int use_beam, int use_pair,
                    double mass_coef, const double* __restrict__ M, KT* __restrict__ K) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= n_blocks) return;
    using A = typename std::conditional<sizeof(KT) == 8, double, float>::type;
    A v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = A(0);
    if (use_beam) {
        for (int q = B.bptr[blk]; q < B.bptr[blk + 1]; ++q) {
            int code = B.bsrc[q];
            int e = code >> 4, lb = code & 15;
#pragma unroll
            for (int k = 0; k < 9; ++k) v[k] += A(Ke[(size_t)(lb * 9 + k) * n_elems + e]);
        }
    }
    if (use_pair && B.pptr) {
        for (int q = B.pptr[blk]; q < B.pptr[blk + 1]; ++q) {
            int code = B.psrc[q];
            int p = code >> 4, k = (code >> 2) & 3, l = code & 3;
            A cc = A(sd[(size_t)k * n_sinc + p] * sd[(size_t)l * n_sinc + p]);
#pragma unroll
            for (int m = 0; m < 9; ++m) v[m] += cc * A(sg[(size_t)m * n_sinc + p]);
        }
    }
    const int gr = B.brow[blk];
    const int gc = B.gcol[blk];
    if (mass_coef != 0.0 && gr == gc) {
#pragma unroll
        for (int a = 0; a < 3; ++a) v[4 * a] += A(mass_coef * M[3 * gr + a]);
    }
    const int r0 = B.gptr[gr];
    const int deg = B.gptr[gr + 1] - r0;
    const size_t base = 9 * (size_t)r0 + 3 * (size_t)(blk - r0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) K[base + (size_t)a * 3 * deg + b] = KT(v[3 * a + b]);
}

template <class T>
__global__ void fill_kernel(size_t n, T* p, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- launchers

hipError_t launch_node_gather(int n_nodes, int n_elems, int n_sinc, const NodeLists& L, const double* fe,
                              const double* sd, int mode, const double* u, const double* f_ext,
                              const uint8_t* fixed, const DynArgs& dy, double* out, double* partial,
                              int* flags, hipStream_t st) {
    if (n_nodes <= 0) return hipSuccess;
    dim3 blk(256), grd((n_nodes + 255) / 256);
    hipLaunchKernelGGL(node_gather_kernel, grd, blk, 0, st, n_nodes, n_elems, n_sinc, L, fe, sd, mode, u,
                       f_ext, fixed, dy, out, partial, flags);
    return hipGetLastError();
}

int node_gather_blocks(int n_nodes) { return (n_nodes + 255) / 256; }

hipError_t launch_sum_partials(int n, const double* partial, double* out, hipStream_t st) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, n, partial, out);
    return hipGetLastError();
}

hipError_t launch_block_gather(int prec, int n_blocks, int n_elems, int n_sinc, const BlockLists& B,
                               const void* Ke, const double* sd, const void* sg, int use_beam, int use_pair,
                               double mass_coef, const double* M, void* K, hipStream_t st) {
    if (n_blocks <= 0) return hipSuccess;
    dim3 blk(256), grd((n_blocks + 255) / 256);
    if (prec == 0)
        hipLaunchKernelGGL((block_gather_kernel<double, double, double>), grd, blk, 0, st, n_blocks, n_elems,
                           n_sinc, B, (const double*)Ke, sd, (const double*)sg, use_beam, use_pair, mass_coef,
                           M, (double*)K);
    else
        hipLaunchKernelGGL((block_gather_kernel<float, float, float>), grd, blk, 0, st, n_blocks, n_elems,
                           n_sinc, B, (const float*)Ke, sd, (const float*)sg, use_beam, use_pair, mass_coef, M,
                           (float*)K);
    return hipGetLastError();
}

hipError_t launch_fill_i32(size_t n, int* p, int v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel<int>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p, v);
    return hipGetLastError();
}

}  // namespace xk
