// Learned candidate prescreen: a small MLP (8 -> 64 -> 64 -> 1, ReLU) that
// chooses which candidate pairs the per-iteration narrowphase watches.
//
// Per pair, features from the increment-start geometry (normalised by
// r_a + r_b): gap now, gap change since the previous increment, centreline
// distance, midpoint lower bound of the segment distance, |cos| of the angle
// between the segments, the two layer ids, and whether the pair contributes.
// Inference runs on the matrix cores (v_mfma_f32_32x32x8_bf16: BF16 operands,
// FP32 accumulation), one 32-pair tile per wave.  Pairs whose logit reaches
// the threshold, and every contributing pair, form the watch list; the rest
// are checked once per Newton call by the exact narrowphase (the verification pass).
//
// Training runs on the GPU too: a per-sample forward / backward on the vector
// ALUs with BF16-rounded weights and inputs, per-workgroup gradient partials
// summed in a fixed order (deterministic), and Adam.  Parameter layout (FP32,
// 4801 values): W1 [8][64], b1 [64], W2 [64][64], b2 [64], w3 [64], b3.

#include <mutex>
#include <string>

#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"
#include "../../include/xkamd.h"

namespace xk {

constexpr int LP_IN = 8, LP_H = 64;
constexpr int OW1 = 0, OB1 = OW1 + LP_IN * LP_H, OW2 = OB1 + LP_H, OB2 = OW2 + LP_H * LP_H, OW3 = OB2 + LP_H,
              OB3 = OW3 + LP_H, LP_NPARAM = OB3 + 1;
static_assert(LP_NPARAM == 4801, "parameter layout");

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t b = __float_as_uint(x);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ float bfr(float x) { return bf2f(f2bf(x)); }

// ------------------------------------------------------------ features
__global__ void lp_features_kernel(int64_t n, PairSoA P, const int* __restrict__ ea, const int* __restrict__ eb,
                                   const double* __restrict__ Xr, const double* __restrict__ u,
                                   const int* __restrict__ elem_layer, float layer_scale,
                                   double* __restrict__ gprev, int64_t n_prev, uint16_t* __restrict__ feat,
                                   float* __restrict__ featf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k[4] = {P.na0[i], P.na1[i], P.nb0[i], P.nb1[i]};
    double x[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) x[q][c] = Xr[3 * k[q] + c] + (u ? u[6 * (size_t)k[q] + c] : 0.0);
    double dA[3], dB[3], dm[3];
    double LA = 0, LB = 0, dmid = 0, ab = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dA[c] = x[1][c] - x[0][c];
        dB[c] = x[3][c] - x[2][c];
        dm[c] = 0.5 * (x[0][c] + x[1][c]) - 0.5 * (x[2][c] + x[3][c]);
        LA += dA[c] * dA[c];
        LB += dB[c] * dB[c];
        dmid += dm[c] * dm[c];
        ab += dA[c] * dB[c];
    }
    LA = sqrt(LA);
    LB = sqrt(LB);
    dmid = sqrt(dmid);
    const double rr = P.ra[i] + P.rb[i];
    const double ir = rr > 1e-30 ? 1.0 / rr : 0.0;
    const double gap = P.gap[i];
    const double dgap = i < n_prev ? gap - gprev[i] : 0.0;
    const double lb = fmax(dmid - 0.5 * LA - 0.5 * LB, 0.0);
    const double cosang = (LA > 1e-30 && LB > 1e-30) ? fabs(ab) / (LA * LB) : 1.0;
    const int la = elem_layer ? elem_layer[ea[i]] : -1, lbid = elem_layer ? elem_layer[eb[i]] : -1;
    float f[8];
    f[0] = (float)fmin(fmax(gap * ir, -1.0), 8.0);
    f[1] = (float)fmin(fmax(dgap * ir, -2.0), 2.0);
    f[2] = (float)fmin(dmid * ir, 16.0);
    f[3] = (float)fmin(lb * ir, 16.0);
    f[4] = (float)cosang;
    f[5] = la >= 0 ? la * layer_scale : 0.0f;
    f[6] = lbid >= 0 ? lbid * layer_scale : 0.0f;
    f[7] = (P.status[i] != ST_INACTIVE || P.p_n[i] > 0.0) ? 1.0f : 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        feat[8 * i + c] = f2bf(f[c]);
        if (featf) featf[8 * i + c] = f[c];
    }
    gprev[i] = gap;
}

// ------------------------------------------------------------ inference (MFMA)
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int LP_WAVES = 4;

template <class FT>
__device__ __forceinline__ short feat_bits(FT v);
template <>
__device__ __forceinline__ short feat_bits<uint16_t>(uint16_t v) { return (short)v; }
template <>
__device__ __forceinline__ short feat_bits<float>(float v) { return (short)f2bf(v); }

// Operand layout of v_mfma_f32_32x32x8_bf16 (wave64): lane l supplies A
// row l % 32, k = 4 (l / 32) + j and B column l % 32, k = 4 (l / 32) + j
// (j = 0..3); result register r of lane l is row 8 (r / 4) + 4 (l / 32) + r % 4,
// column l % 32.
template <class FT>
__global__ void __launch_bounds__(64 * LP_WAVES)
lp_infer_kernel(int64_t n, const FT* __restrict__ feat, const float* __restrict__ prm, float thr,
                const int* __restrict__ sinc, int64_t* __restrict__ keep, float* __restrict__ logits) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    __shared__ uint16_t H1[LP_WAVES][32][LP_H + 8];
    __shared__ float S[LP_WAVES][32][33];
    bf16x4 bw1[2], bw2[2][8];
    float b1v[2], b2v[2], w3v[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        const int col = 32 * nb + lr;
#pragma unroll
        for (int j = 0; j < 4; ++j) bw1[nb][j] = (short)f2bf(prm[OW1 + (4 * lh + j) * LP_H + col]);
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int j = 0; j < 4; ++j) bw2[nb][s][j] = (short)f2bf(prm[OW2 + (8 * s + 4 * lh + j) * LP_H + col]);
        b1v[nb] = prm[OB1 + col];
        b2v[nb] = prm[OB2 + col];
        w3v[nb] = prm[OW3 + col];
    }
    const float b3 = prm[OB3];
    const int64_t n_tiles = (n + 31) / 32;
    // the tile loop is uniform over the workgroup (barriers inside)
    for (int64_t base = (int64_t)blockIdx.x * LP_WAVES; base < n_tiles; base += (int64_t)gridDim.x * LP_WAVES) {
        const int64_t r0 = (base + wave) * 32;
        bf16x4 ax;
        {
            const int64_t row = r0 + lr;
#pragma unroll
            for (int j = 0; j < 4; ++j) ax[j] = row < n ? feat_bits<FT>(feat[8 * row + 4 * lh + j]) : (short)0;
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f32x16 c = {};
            c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ax, bw1[nb], c, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float h = fmaxf(c[r] + b1v[nb], 0.0f);
                H1[wave][8 * (r >> 2) + 4 * lh + (r & 3)][32 * nb + lr] = f2bf(h);
            }
        }
        __syncthreads();
        float part[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) part[r] = 0.0f;
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f32x16 c = {};
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                bf16x4 a;
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = (short)H1[wave][lr][8 * s + 4 * lh + j];
                c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bw2[nb][s], c, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) part[r] += fmaxf(c[r] + b2v[nb], 0.0f) * w3v[nb];
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) S[wave][8 * (r >> 2) + 4 * lh + (r & 3)][lr] = part[r];
        __syncthreads();
        if (lane < 32) {
            const int64_t row = r0 + lane;
            float z = b3;
            for (int c = 0; c < 32; ++c) z += S[wave][lane][c];
            if (row < n) {
                if (logits) logits[row] = z;
                if (keep) keep[row] = (z >= thr || (sinc && sinc[row] >= 0)) ? 1 : 0;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ list split
// keep flags + their exclusive scan -> watch list (kept, pair order) and
// prune list (the rest, pair order).
__global__ void lp_split_kernel(int64_t n, const int64_t* __restrict__ keep, const int64_t* __restrict__ pos,
                                int* __restrict__ watch, int* __restrict__ prune) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keep[i]) watch[pos[i]] = (int)i;
    else prune[i - pos[i]] = (int)i;
}

// #(gap <= g_on) over the listed pairs (integer count: order-free)
__global__ void lp_miss_kernel(int64_t n, const int* __restrict__ idx, const double* __restrict__ gap, double g_on,
                               int* __restrict__ count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int m = 0;
    if (t < n) m = gap[idx[t]] <= g_on ? 1 : 0;
    const unsigned long long b = __ballot(m);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, __popcll(b));
}

// ------------------------------------------------------------ compaction
// gather (scatter == 0): b[c][j] = a[c][idx[j]]; scatter: b[c][idx[j]] = a[c][j]
template <class T>
__global__ void permute_kernel(int64_t n, int comps, const T* __restrict__ a, int64_t sa, T* __restrict__ b,
                               int64_t sb, const int* __restrict__ idx, int scatter) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t k = idx[j];
    for (int c = 0; c < comps; ++c) {
        if (scatter) b[c * sb + k] = a[c * sa + j];
        else b[c * sb + j] = a[c * sa + k];
    }
}

// Several arrays of one permutation in one launch: blockIdx.y picks the
// array (element size and component count per array), each lane moves one
// element's components as permute_kernel does.
__global__ void permute_set_kernel(int64_t n, PermuteSet s, int64_t sa, int64_t sb, const int* __restrict__ idx,
                                   int scatter) {
    const int q = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t k = idx[j];
    const int64_t src = scatter ? j : k, dst = scatter ? k : j;
    const int comps = s.comps[q];
    switch (s.esz[q]) {   // uniform over the workgroup
        case 1: {
            const int8_t* a = (const int8_t*)s.a[q];
            int8_t* b = (int8_t*)s.b[q];
            for (int c = 0; c < comps; ++c) b[c * sb + dst] = a[c * sa + src];
            break;
        }
        case 4: {
            const int* a = (const int*)s.a[q];
            int* b = (int*)s.b[q];
            for (int c = 0; c < comps; ++c) b[c * sb + dst] = a[c * sa + src];
            break;
        }
        default: {
            const double* a = (const double*)s.a[q];
            double* b = (double*)s.b[q];
            for (int c = 0; c < comps; ++c) b[c * sb + dst] = a[c * sa + src];
            break;
        }
    }
}

hipError_t launch_permute_set(const PermuteSet& s, int64_t n, int64_t sa, int64_t sb, const int* idx, int scatter,
                              hipStream_t st) {
    if (n <= 0 || s.count <= 0) return hipSuccess;
    for (int q = 0; q < s.count; ++q)
        if (s.esz[q] != 1 && s.esz[q] != 4 && s.esz[q] != 8) return hipErrorInvalidValue;
    const dim3 g((unsigned)((n + 255) / 256), (unsigned)s.count), blk(256);
    hipLaunchKernelGGL(permute_set_kernel, g, blk, 0, st, n, s, sa, sb, idx, scatter);
    return hipGetLastError();
}

hipError_t launch_permute(int esz, int64_t n, int comps, const void* a, int64_t sa, void* b, int64_t sb,
                          const int* idx, int scatter, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const dim3 g((unsigned)((n + 255) / 256)), blk(256);
    switch (esz) {
        case 1: hipLaunchKernelGGL(permute_kernel<int8_t>, g, blk, 0, st, n, comps, (const int8_t*)a, sa, (int8_t*)b,
                                   sb, idx, scatter); break;
        case 4: hipLaunchKernelGGL(permute_kernel<int>, g, blk, 0, st, n, comps, (const int*)a, sa, (int*)b, sb, idx,
                                   scatter); break;
        case 8: hipLaunchKernelGGL(permute_kernel<double>, g, blk, 0, st, n, comps, (const double*)a, sa, (double*)b,
                                   sb, idx, scatter); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ training
constexpr int TG = 64;   // samples per chunk = lanes per workgroup
constexpr int TP = LP_H + 1;

__global__ void __launch_bounds__(TG)
lp_grad_kernel(int64_t n, const float* __restrict__ X, const float* __restrict__ y, const float* __restrict__ prm,
               float pos_w, float inv_n, float* __restrict__ partial) {
    __shared__ float W1s[LP_IN][LP_H], W2s[LP_H][LP_H], b1s[LP_H], b2s[LP_H], w3s[LP_H];
    __shared__ float Xs[TG][LP_IN], A1[TG][TP], A2[TG][TP], D1[TG][TP], D2[TG][TP], G3[TG], Ls[TG];
    const int t = threadIdx.x;
    for (int q = t; q < LP_IN * LP_H; q += TG) W1s[q / LP_H][q % LP_H] = bfr(prm[OW1 + q]);
    for (int q = t; q < LP_H * LP_H; q += TG) W2s[q / LP_H][q % LP_H] = bfr(prm[OW2 + q]);
    b1s[t] = prm[OB1 + t];
    b2s[t] = prm[OB2 + t];
    w3s[t] = prm[OW3 + t];
    const float b3 = prm[OB3];
    float gW2[LP_H], gW1[LP_IN], gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
#pragma unroll
    for (int j = 0; j < LP_H; ++j) gW2[j] = 0.f;
#pragma unroll
    for (int k = 0; k < LP_IN; ++k) gW1[k] = 0.f;
    __syncthreads();
    const int64_t n_chunks = (n + TG - 1) / TG;
    for (int64_t ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
        const int64_t s = ch * TG + t;
        const bool live = s < n;
        // ---- per-sample forward / backward (lane t = sample s)
        float x[LP_IN];
#pragma unroll
        for (int k = 0; k < LP_IN; ++k) {
            x[k] = live ? bfr(X[LP_IN * s + k]) : 0.f;
            Xs[t][k] = x[k];
        }
        for (int j = 0; j < LP_H; ++j) {
            float z = b1s[j];
#pragma unroll
            for (int k = 0; k < LP_IN; ++k) z += x[k] * W1s[k][j];
            A1[t][j] = z;   // pre-activation for now
        }
        float z2[LP_H];
#pragma unroll
        for (int j = 0; j < LP_H; ++j) z2[j] = b2s[j];
        for (int i = 0; i < LP_H; ++i) {
            const float a = fmaxf(A1[t][i], 0.f);
#pragma unroll
            for (int j = 0; j < LP_H; ++j) z2[j] += a * W2s[i][j];
        }
        float z3 = b3;
#pragma unroll
        for (int j = 0; j < LP_H; ++j) {
            const float a2 = fmaxf(z2[j], 0.f);
            A2[t][j] = a2;
            z3 += a2 * w3s[j];
        }
        const float yy = live ? y[s] : 0.f;
        const float sp = z3 > 0.f ? z3 + log1pf(expf(-z3)) : log1pf(expf(z3));    // softplus(z3)
        const float spm = sp - z3;                                                // softplus(-z3)
        const float sig = 0.5f * (1.f + tanhf(0.5f * z3));
        const float g3 = live ? (-pos_w * yy * (1.f - sig) + (1.f - yy) * sig) * inv_n : 0.f;
        Ls[t] = live ? pos_w * yy * spm + (1.f - yy) * sp : 0.f;
        G3[t] = g3;
        float d2[LP_H];
#pragma unroll
        for (int j = 0; j < LP_H; ++j) {
            d2[j] = z2[j] > 0.f ? g3 * w3s[j] : 0.f;
            D2[t][j] = d2[j];
        }
        for (int i = 0; i < LP_H; ++i) {
            float d = 0.f;
#pragma unroll
            for (int j = 0; j < LP_H; ++j) d += W2s[i][j] * d2[j];
            const float z1 = A1[t][i];
            D1[t][i] = z1 > 0.f ? d : 0.f;
            A1[t][i] = fmaxf(z1, 0.f);
        }
        __syncthreads();
        // ---- gradient accumulation over the chunk, samples in order (lane t = column t)
        const int m = (int)min<int64_t>(TG, n - ch * TG);
        for (int q = 0; q < m; ++q) {
            const float a1t = A1[q][t], d2t = D2[q][t], d1t = D1[q][t];
#pragma unroll
            for (int j = 0; j < LP_H; ++j) gW2[j] += a1t * D2[q][j];
#pragma unroll
            for (int k = 0; k < LP_IN; ++k) gW1[k] += Xs[q][k] * d1t;
            gb1 += d1t;
            gb2 += d2t;
            gw3 += G3[q] * A2[q][t];
            if (t == 0) {
                gb3 += G3[q];
                loss += Ls[q];
            }
        }
        __syncthreads();
    }
    float* out = partial + (size_t)blockIdx.x * (LP_NPARAM + 1);
    // gW2[j] is dL/dW2[t][j]; gW1[k] is dL/dW1[k][t]
#pragma unroll
    for (int j = 0; j < LP_H; ++j) out[OW2 + t * LP_H + j] = gW2[j];
#pragma unroll
    for (int k = 0; k < LP_IN; ++k) out[OW1 + k * LP_H + t] = gW1[k];
    out[OB1 + t] = gb1;
    out[OB2 + t] = gb2;
    out[OW3 + t] = gw3;
    if (t == 0) {
        out[OB3] = gb3;
        out[LP_NPARAM] = loss * inv_n;
    }
}

__global__ void lp_reduce_kernel(int nb, const float* __restrict__ partial, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > LP_NPARAM) return;
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += partial[(size_t)b * (LP_NPARAM + 1) + p];
    grad[p] = s;
}

__global__ void lp_adam_kernel(float* __restrict__ prm, const float* __restrict__ g, float* __restrict__ m,
                               float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= LP_NPARAM) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    prm[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// ------------------------------------------------------------ assembly history
// Each fused assembly over the watch list appends the translational
// displacement it was given (3 values per node) to slot *count of hist, if
// the slot exists; a one-lane kernel then advances *count.  The verification
// pass replays these geometries for the pruned pairs.
__global__ void lp_hist_push_kernel(int64_t n3, const double* __restrict__ u, double* __restrict__ hist,
                                    int64_t stride, int cap, const int* __restrict__ count,
                                    const double* __restrict__ stop) {
    if (!nc_go(stop)) return;   // no assembly ran: nothing to record
    const int slot = *count;
    if (slot >= cap) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const int64_t node = q / 3, c = q - 3 * node;
    hist[(size_t)slot * stride + q] = u[6 * node + c];
}

__global__ void lp_hist_bump_kernel(int* __restrict__ count, const double* __restrict__ stop) {
    if (threadIdx.x == 0 && nc_go(stop)) *count += 1;
}

// ------------------------------------------------------------ launchers
int lp_n_params() { return LP_NPARAM; }

int lp_grad_blocks(int64_t n) { return (int)std::min<int64_t>((n + TG - 1) / TG, 256); }

hipError_t launch_lp_features(int64_t n, const PairSoA& P, const int* ea, const int* eb, const double* Xr,
                              const double* u, const int* elem_layer, float layer_scale, double* gprev,
                              int64_t n_prev, uint16_t* feat, float* featf, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lp_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, P, ea, eb, Xr, u,
                       elem_layer, layer_scale, gprev, n_prev, feat, featf);
    return hipGetLastError();
}

static unsigned infer_grid(int64_t n) {
    const int64_t tiles = (n + 31) / 32;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + LP_WAVES - 1) / LP_WAVES, 2048));
}

hipError_t launch_lp_infer_bf16(int64_t n, const uint16_t* feat, const float* prm, float thr, const int* sinc,
                                int64_t* keep, float* logits, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lp_infer_kernel<uint16_t>, dim3(infer_grid(n)), dim3(64 * LP_WAVES), 0, st, n, feat, prm,
                       thr, sinc, keep, logits);
    return hipGetLastError();
}

hipError_t launch_lp_infer_f32(int64_t n, const float* feat, const float* prm, float thr, int64_t* keep,
                               float* logits, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lp_infer_kernel<float>, dim3(infer_grid(n)), dim3(64 * LP_WAVES), 0, st, n, feat, prm, thr,
                       (const int*)nullptr, keep, logits);
    return hipGetLastError();
}

hipError_t launch_lp_split(int64_t n, const int64_t* keep, const int64_t* pos, int* watch, int* prune,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lp_split_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, keep, pos, watch,
                       prune);
    return hipGetLastError();
}

hipError_t launch_lp_miss(int64_t n, const int* idx, const double* gap, double g_on, int* count, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lp_miss_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, idx, gap, g_on, count);
    return hipGetLastError();
}

hipError_t launch_lp_hist_push(int n_nodes, const double* u, double* hist, int cap, int* count, hipStream_t st,
                               const double* stop) {
    const int64_t n3 = 3 * (int64_t)n_nodes;
    if (n3 <= 0) return hipSuccess;
    hipLaunchKernelGGL(lp_hist_push_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, n3, u, hist, n3,
                       cap, (const int*)count, stop);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(lp_hist_bump_kernel, dim3(1), dim3(64), 0, st, count, stop);
    return hipGetLastError();
}

hipError_t launch_lp_grad(int64_t n, const float* X, const float* y, const float* prm, float pos_w, float* partial,
                          float* grad, hipStream_t st) {
    const int nb = lp_grad_blocks(n);
    if (n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lp_grad_kernel, dim3(nb), dim3(TG), 0, st, n, X, y, prm, pos_w, (float)(1.0 / (double)n),
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(lp_reduce_kernel, dim3((LP_NPARAM + 1 + 255) / 256), dim3(256), 0, st, nb, partial, grad);
    return hipGetLastError();
}

hipError_t launch_lp_adam(float* prm, const float* g, float* m, float* v, int t, float lr, float b1, float b2,
                          float eps, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(lp_adam_kernel, dim3((LP_NPARAM + 255) / 256), dim3(256), 0, st, prm, g, m, v, lr, b1, b2, eps,
                       c1, c2);
    return hipGetLastError();
}

// ------------------------------------------------------------ allocation-free gradient
// The per-workgroup gradient partials of xk_mlp_grad_noalloc live in this
// fixed array instead of a caller workspace: one row of LP_NPARAM + 1 floats
// (gradient, then loss) per workgroup of lp_grad_kernel, whose grid never
// exceeds 256 workgroups (lp_grad_blocks), so 256 x 4802 floats = 4.9 MB.
// Every GPU the library runs on holds its own copy until the process ends.
// Host code never passes `scratch` itself to a kernel: its host-side name is
// not a device address.  The device address is looked up once per device.
constexpr int64_t LP_SCRATCH_ROWS = 256;
constexpr int64_t LP_SCRATCH_FLOATS = LP_SCRATCH_ROWS * (LP_NPARAM + 1);
static __device__ __align__(256) float scratch[LP_SCRATCH_FLOATS];

constexpr int LP_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_dev[LP_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (resolved on the first
// call on that device, with that device current).
static int lp_scratch_ptr(float** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("xk_mlp_grad_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return -1;
    }
    if (dev < 0 || dev >= LP_MAX_DEVICES) {
        set_error("xk_mlp_grad_noalloc: device ordinal out of range");
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("xk_mlp_grad_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return -1;
        }
        g_scratch_dev[dev] = (float*)p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

}  // namespace xk

extern "C" int xk_mlp_grad_noalloc(const float* X, const float* y, int64_t n, const float* params,
                                   float pos_weight, float* grad_and_loss, void* stream) {
    if (n <= 0 || !X || !y || !params || !grad_and_loss) {
        xk::set_error("xk_mlp_grad_noalloc: invalid arguments");
        return -1;
    }
    const int64_t need = (int64_t)xk::lp_grad_blocks(n) * (xk::LP_NPARAM + 1);
    if (need > xk::LP_SCRATCH_FLOATS) {
        xk::set_error("xk_mlp_grad_noalloc: " + std::to_string(n) + " rows need " + std::to_string(need) +
                      " workspace floats, more than the fixed " + std::to_string(xk::LP_SCRATCH_FLOATS));
        return -1;
    }
    float* work = nullptr;
    if (xk::lp_scratch_ptr(&work) != 0) return -1;
    // every partial row the reduction reads is written by this call's gradient
    // kernel, so nothing needs clearing between calls
    hipError_t e = xk::launch_lp_grad(n, X, y, params, pos_weight, work, grad_and_loss, (hipStream_t)stream);
    if (e != hipSuccess) {
        xk::set_error(std::string("xk_mlp_grad_noalloc: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}
