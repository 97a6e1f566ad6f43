// Deterministic gathers of staged element / pair contributions into the
// global residual and the block-CSR tangent, plus the Generalized-alpha
// effective terms, the FP64 residual norm and the active-pair count -- one
// launch.
//
// Every output entry is summed by one lane over a contribution list fixed at
// pattern-build time (elements in element order, then pairs in pair order), so
// results are bitwise reproducible run to run (no floating-point atomics).
// Lanes [0, 6 n_nodes) own one residual DOF each.  The following lanes own
// tangent entries: on the direct path (element kernel already wrote the beam
// blocks) one lane per 3x3 block with pair contributions, driven by a 16-byte
// descriptor {CSR offset, row stride, list begin, list end | beam bit}; on the
// staged path one row of one 3x3 block of the whole pattern.  The residual
// partial sums and the pair kernel's per-workgroup active counts are reduced by
// a one-workgroup finish kernel (no atomics, no memset, capturable in a HIP
// graph).
//
// Residual follows ContactForceAssemblyProcess
// (xkep_cae/contact/solver/_newton_steps.py:124-129) and the dynamic steps
// of _newton_dynamic.py:189-192 with GeneralizedAlphaProcess.correct /
// effective_residual (time_integration/strategy.py:178-216):
//   R = f_int - (f_c + f_fric) - f_ext ; R[fixed] = 0
//   acc = c0 (u - u_pred) ; vel = v_pred + dt gamma acc
//   R += M ((1 - alpha_m) acc + alpha_m acc_old) ; R[fixed] = 0
// Tangent follows TangentAssemblyProcess (:289-313) and effective_stiffness
// (time_integration/strategy.py:190-203):
//   K = K_beam + K_c - K_fric + (1 - alpha_m) c0 M.

#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"
#include <type_traits>

namespace xk {

// Sum over the 64 lanes of a wave by a butterfly: every lane gets the same
// value, formed in the same order each call.
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// node_mode: -1 no residual work, 0 = f_int only, 1 = pair force only, 2 = full residual
template <class KT, class ST, class GT, int GB>
__global__ void __launch_bounds__(GB) gather_kernel(GatherArgs g, const ST* __restrict__ Ke,
                                                     const GT* __restrict__ sg, KT* __restrict__ K) {
    if (!nc_go(g.stop)) return;   // the Newton call this attempt belongs to has ended
    __shared__ double red[GB / 64];
    double ss = 0.0;
    const bool node_block = (int)blockIdx.x < g.node_blocks;
    if (node_block) {
        const int tid = blockIdx.x * GB + threadIdx.x;
        if (tid < 6 * g.n_nodes) {
            const int i = tid / 6, d = tid - 6 * (tid / 6);
            double fb = 0.0, fp = 0.0;
            if (g.node_mode != 1 && g.fint) {
                fb = g.fint[tid];
            } else if (g.node_mode != 1) {
                const int q0 = g.L.bptr[i], q1 = g.L.bptr[i + 1];
                for (int q = q0; q < q1; ++q) {
                    const int code = g.L.bsrc[q];
                    fb += g.fe[12 * (size_t)(code >> 1) + 6 * (code & 1) + d];
                }
            }
            if (g.node_mode != 0 && d < 3 && g.L.pptr) {
                // four loads in flight per step, summed in list order
                const int q0 = g.L.pptr[i], q1 = g.L.pptr[i + 1];
                int q = q0;
                for (; q + 4 <= q1; q += 4) {
                    const double a0 = g.nf[3 * (size_t)q + d], a1 = g.nf[3 * (size_t)(q + 1) + d];
                    const double a2 = g.nf[3 * (size_t)(q + 2) + d], a3 = g.nf[3 * (size_t)(q + 3) + d];
                    fp = (((fp + a0) + a1) + a2) + a3;
                }
                for (; q < q1; ++q) fp += g.nf[3 * (size_t)q + d];
            }
            double r;
            if (g.node_mode == 0) {
                r = fb;
            } else if (g.node_mode == 1) {
                r = d < 3 ? fp : 0.0;
            } else {
                r = (fb + (d < 3 ? -fp : -0.0)) - g.f_ext[tid];
                const bool fx = g.fixed[tid] != 0;
                if (fx) r = 0.0;
                if (g.dy.on) {
                    const double ac = g.dy.c0 * (g.u[tid] - g.dy.u_pred[tid]);
                    // acc and vel are state for the increment's end, not read by the
                    // following kernels: stored non-temporally (cache kept for the inputs),
                    // and not at all when the caller recomputes them there (null acc)
                    if (g.dy.acc) {
                        __builtin_nontemporal_store(ac, &g.dy.acc[tid]);
                        __builtin_nontemporal_store(g.dy.v_pred[tid] + g.dy.dt * g.dy.gamma * ac, &g.dy.vel[tid]);
                    }
                    const double amid = (1.0 - g.dy.alpha_m) * ac + g.dy.alpha_m * g.dy.acc_old[tid];
                    r = r + g.dy.M[tid] * amid;
                    if (fx) r = 0.0;
                }
                ss = r * r;
                if (!isfinite(r)) atomicOr(g.flags, FLAG_NONFINITE);
            }
            g.out[tid] = r;
        }
    } else if (K && g.direct) {
        // direct path: one lane per block with pair contributions
        const int kk = (blockIdx.x - g.node_blocks) * GB + threadIdx.x;
        const int4 dsc = kk < g.n_cdesc ? g.cdesc[kk] : int4{0, 0, 0, 0};
        const bool beam = dsc.w < 0 && g.use_beam;
        // without pair terms a beam block already holds its final value;
        // contact-only blocks are written (zero) either way
        if (kk < g.n_cdesc && (g.use_pair || !beam)) {
            using A = typename std::conditional<sizeof(KT) == 8, double, float>::type;
            const size_t base = (size_t)(unsigned)dsc.x;
            const size_t rs = (size_t)dsc.y;
            const int q0 = dsc.z, q1 = g.use_pair ? (dsc.w & 0x7fffffff) : q0;
            A v[9];
            if (beam) {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) v[3 * a + b] = A(K[base + a * rs + b]);
            } else {
#pragma unroll
                for (int k = 0; k < 9; ++k) v[k] = A(0);
            }
            // two contributions per step: both codes, then both records, in
            // flight together; accumulated in list order
            int q = q0;
            for (; q + 2 <= q1; q += 2) {
                const int c0 = g.B.psrc[q], c1 = g.B.psrc[q + 1];
                const GT* r0 = sg + SG_REC * (size_t)(c0 >> 4);
                const GT* r1 = sg + SG_REC * (size_t)(c1 >> 4);
                GT x0[9], x1[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    x0[k] = r0[k];
                    x1[k] = r1[k];
                }
                const A cc0 = A(r0[SG_C + ((c0 >> 2) & 3)]) * A(r0[SG_C + (c0 & 3)]);
                const A cc1 = A(r1[SG_C + ((c1 >> 2) & 3)]) * A(r1[SG_C + (c1 & 3)]);
#pragma unroll
                for (int k = 0; k < 9; ++k) v[k] += cc0 * A(x0[k]);
#pragma unroll
                for (int k = 0; k < 9; ++k) v[k] += cc1 * A(x1[k]);
            }
            if (q < q1) {
                const int code = g.B.psrc[q];
                const GT* r = sg + SG_REC * (size_t)(code >> 4);
                const A cc = A(r[SG_C + ((code >> 2) & 3)]) * A(r[SG_C + (code & 3)]);
#pragma unroll
                for (int k = 0; k < 9; ++k) v[k] += cc * A(r[k]);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) K[base + a * rs + b] = KT(v[3 * a + b]);
        }
    } else if (K) {
        // staged path: one lane per row of every block of the pattern
        const int tid = (blockIdx.x - g.node_blocks) * GB + threadIdx.x;
        if (tid < 3 * g.n_blocks) {
            const int blk = tid / 3, a = tid - 3 * (tid / 3);
            using A = typename std::conditional<sizeof(KT) == 8, double, float>::type;
            A v0 = A(0), v1 = A(0), v2 = A(0);
            const int gr0 = g.B.brow[blk];
            const int r00 = g.B.gptr[gr0];
            const int deg0 = g.B.gptr[gr0 + 1] - r00;
            const size_t base0 = 9 * (size_t)r00 + 3 * (size_t)(blk - r00) + (size_t)a * 3 * deg0;
            if (g.use_beam) {
                const int q0 = g.B.bptr[blk], q1 = g.B.bptr[blk + 1];
                for (int q = q0; q < q1; ++q) {
                    const int code = g.B.bsrc[q];
                    // element record: 144 values, 3x3 group block (code & 15) at 9 (code & 15)
                    const ST* r = Ke + 144 * (size_t)(code >> 4) + 9 * (code & 15) + 3 * a;
                    v0 += A(r[0]);
                    v1 += A(r[1]);
                    v2 += A(r[2]);
                }
            }
            if (g.use_pair && g.B.pptr) {
                const int q0 = g.B.pptr[blk], q1 = g.B.pptr[blk + 1];
                for (int q = q0; q < q1; ++q) {
                    const int code = g.B.psrc[q];
                    const GT* r = sg + SG_REC * (size_t)(code >> 4);
                    const A cc = A(r[SG_C + ((code >> 2) & 3)]) * A(r[SG_C + (code & 3)]);
                    v0 += cc * A(r[3 * a]);
                    v1 += cc * A(r[3 * a + 1]);
                    v2 += cc * A(r[3 * a + 2]);
                }
            }
            if (g.mass_coef != 0.0 && gr0 == g.B.gcol[blk]) {
                const A m = A(g.mass_coef * g.M[3 * gr0 + a]);
                if (a == 0) v0 += m;
                else if (a == 1) v1 += m;
                else v2 += m;
            }
            K[base0] = KT(v0);
            K[base0 + 1] = KT(v1);
            K[base0 + 2] = KT(v2);
        }
    }
    // residual partial of this workgroup: butterfly within each wave, then the
    // waves' sums in wave order (fixed shape: deterministic)
    if (g.partial && node_block) {
        ss = wave_sum(ss);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = red[0];
            for (int w = 1; w < GB / 64; ++w) t += red[w];
            g.partial[blockIdx.x] = t;
        }
    }
}

// ||R||^2 from the gather's workgroup partials and #(p_n > 0) from the pair
// kernel's workgroup counts, one workgroup, fixed summation order.  (A
// last-workgroup reduction inside the gather would need a device-scope release
// fence per workgroup, which on a multi-XCD part writes back L2 each time.)
constexpr int FB = 1024;  // finish kernel lanes

__global__ void __launch_bounds__(FB) finish_kernel(int n_part, const double* __restrict__ partial,
                                                    int n_act, const int* __restrict__ act_part,
                                                    double* __restrict__ res_norm2, int* __restrict__ n_active,
                                                    const double* __restrict__ stop) {
    if (!nc_go(stop)) return;
    __shared__ double red[FB / 64];
    __shared__ int ired[FB / 64];
    // each lane sums its strided entries in index order; one chunk issues all its
    // loads (8 partials and 16 counts per lane) before the first add
    constexpr int CP = 8, CA = 16;
    double s = 0.0;
    int a = 0;
    for (int k0 = threadIdx.x; k0 < n_part || k0 < n_act; k0 += CA * FB) {
        double p[CP];
        int q[CA];
#pragma unroll
        for (int j = 0; j < CP; ++j) {
            const int k = k0 + j * FB;
            p[j] = k < n_part ? partial[k] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            const int k = k0 + j * FB;
            q[j] = k < n_act ? act_part[k] : 0;
        }
#pragma unroll
        for (int j = 0; j < CP; ++j) s += p[j];
#pragma unroll
        for (int j = CP; j < CA; ++j) {
            const int k = k0 + j * FB;
            if (k < n_part) s += partial[k];
        }
#pragma unroll
        for (int j = 0; j < CA; ++j) a += q[j];
    }
    s = wave_sum(s);
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = s;
        ired[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        int c = ired[0];
        for (int w = 1; w < FB / 64; ++w) {
            t += red[w];
            c += ired[w];
        }
        if (res_norm2) *res_norm2 = t;
        if (n_active) *n_active = c;
    }
}

template <class T>
__global__ void fill_kernel(size_t n, T* p, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- launchers

int gather_node_blocks(int n_nodes, int bs) { return (6 * n_nodes + bs - 1) / bs; }

bool gather_block_ok(int bs) { return bs == 128 || bs == 256 || bs == 512; }

template <class KT, class GT, int BS>
static void launch_gather_bs(int nb, const GatherArgs& g, const void* Ke, const void* sg, void* K, hipStream_t st) {
    hipLaunchKernelGGL((gather_kernel<KT, KT, GT, BS>), dim3(nb), dim3(BS), 0, st, g, (const KT*)Ke, (const GT*)sg,
                       (KT*)K);
}

// K of type T; the pairs' tangent records are FP64 in both forms
template <class T>
static hipError_t launch_gather_t(int nb, const GatherArgs& g, const void* Ke, const void* sg, void* K,
                                  hipStream_t st) {
    switch (g.block) {
        case 128: launch_gather_bs<T, double, 128>(nb, g, Ke, sg, K, st); break;
        case 256: launch_gather_bs<T, double, 256>(nb, g, Ke, sg, K, st); break;
        case 512: launch_gather_bs<T, double, 512>(nb, g, Ke, sg, K, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gather(int prec, GatherArgs g, const void* Ke, const void* sg, void* K, hipStream_t st,
                         hipEvent_t between) {
    if (K && g.use_pair && !sg) return hipErrorInvalidValue;  // pair terms need the tangent records
    if (!g.block) g.block = 256;
    g.node_blocks = g.node_mode >= 0 ? gather_node_blocks(g.n_nodes, g.block) : 0;
    const int kb = K ? ((g.direct ? g.n_cdesc : 3 * g.n_blocks) + g.block - 1) / g.block : 0;
    const int nb = g.node_blocks + kb;
    if (nb <= 0) return hipSuccess;
    hipError_t e = prec == 0 ? launch_gather_t<double>(nb, g, Ke, sg, K, st) : launch_gather_t<float>(nb, g, Ke, sg, K, st);
    if (e != hipSuccess) return e;
    if (between && (e = hipEventRecord(between, st)) != hipSuccess) return e;
    if (g.partial && (g.res_norm2 || g.n_active))
        hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(FB), 0, st, g.node_blocks, (const double*)g.partial,
                           g.n_act_part, g.act_part, g.res_norm2, g.n_active, g.stop);
    return hipGetLastError();
}

template <class T>
static hipError_t launch_fill(size_t n, T* p, T v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p, v);
    return hipGetLastError();
}

hipError_t launch_fill_i32(size_t n, int* p, int v, hipStream_t st) { return launch_fill<int>(n, p, v, st); }
hipError_t launch_fill_f64(size_t n, double* p, double v, hipStream_t st) { return launch_fill<double>(n, p, v, st); }
hipError_t launch_fill_f32(size_t n, float* p, float v, hipStream_t st) { return launch_fill<float>(n, p, v, st); }

}  // namespace xk
