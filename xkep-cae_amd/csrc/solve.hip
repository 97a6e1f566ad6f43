// Linear solve of one Newton attempt on the device (direct element path),
// the role of LinearSolveProcess (contact/solver/_newton_steps.py:338-366):
//
//   (D K D + I - D) du = -D R            D = diag(free dofs)
//
// The tangent splits as K = A_b + sum_p U_p G_p U_p^T:
//   A_b  = K_beam + (1 - alpha_m) c0 M, block tridiagonal along every element
//          chain (6x6 node blocks; chains never couple through elements);
//   U_p  = c_p (x) I3 on the translational dofs of the pair's four nodes and
//   G_p    the pair's 3x3 tangent block: exactly the pair kernel's staged
//          tangent record, which the gather scatters as c_k c_l G_p.
// With the mask on both (A~_b, U~), the Woodbury identity gives
//   A^-1 r = A~_b^-1 (r - U~ w),   (I + G U~^T A~_b^-1 U~) w = G U~^T A~_b^-1 r.
// A~_b is factored by block LU along each chain (6x6 pivot blocks inverted by
// Gauss-Jordan with partial pivoting); the capacitance matrix
// I + G U~^T A~_b^-1 U~ (3 rows per contributing pair) is dense within a
// connected component of the contact graph (pairs sharing a chain) and is
// LU-factored per component: blocked, without row interchanges, the trailing
// updates by rocBLAS DGEMM (rocSOLVER's partially pivoted getrf when that
// meets a zero pivot or its refinement stalls above the tolerance); the
// triangular solves run in blocks of TB rows, one launch per block.  Penalty springs that tie the soft,
// mass-dominated modes of whole wires make the capacitance ill-conditioned
// (condition numbers near 1e12 on strands): the result's residual is
// U (S w - t), so it carries the capacitance solve's backward error times
// ||S|| ||w|| (relative residuals 1e-4 .. 1e-2 on strands).  Iterative
// refinement against the assembled K contracts it by that factor per step
// and reaches the residual of a sparse direct solve in a few steps.

#include "../../include/xkamd.h"
#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"

#include <algorithm>
#include <cfloat>
#include <numeric>
#include <rocsolver/rocsolver.h>

namespace xk {

enum SolveBuf {
    SV_KB, SV_AB, SV_LF, SV_SINV, SV_SLIST, SV_SN, SV_SG, SV_SC, SV_POS, SV_CSLOT, SV_CN, SV_CNODE, SV_CHM,
    SV_CHCN0, SV_CHYOFF, SV_CHCOFF, SV_Y, SV_C, SV_S, SV_IPIV, SV_INFO, SV_W, SV_B, SV_X, SV_R, SV_T, SV_RED,
    SV_PERM, SV_W2, SV_PIV32, SV_COUNT
};

// ------------------------------------------------------------------ kernels
//
// Every kernel takes the device Newton control state word first (`stop`,
// null outside a device-controlled call) and does nothing once the call it
// belongs to has ended: a solve enqueued past the end of a call is a no-op.

// A double of another lane of the wave (lane uniform)
__device__ __forceinline__ double sv_readlane(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// 32-bit pivots and info of one LU widened into the 64-bit buffers
__global__ void sv_widen_kernel(const double* __restrict__ stop,
                                int64_t n, const int* __restrict__ piv32, int64_t* __restrict__ piv64,
                                const int* __restrict__ info32, int64_t* __restrict__ info64) {
    if (!nc_go(stop)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) piv64[i] = piv32[i];
    if (i == 0) *info64 = *info32;
}

__device__ __forceinline__ bool sv_link(const int8_t* nflag, int n, int n_nodes) {
    return n >= 0 && n < n_nodes && (nflag[n] & 2);
}

__device__ __forceinline__ int sv_find(const int* __restrict__ gptr, const int* __restrict__ gcol, int row, int col) {
    int lo = gptr[row], hi = gptr[row + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (gcol[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Per node n, its 6 scalar rows over the columns of nodes n-1, n, n+1
// (6 x 18, FP64), masked: a fixed row or column becomes the identity's.
template <class KT>
__global__ void sv_extract_kernel(const double* __restrict__ stop,
                                  int n_nodes, const int8_t* __restrict__ nflag, const int* __restrict__ gptr,
                                  const int* __restrict__ gcol, const KT* __restrict__ Kb,
                                  const uint8_t* __restrict__ fixed, double* __restrict__ Ab) {
    if (!nc_go(stop)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * n_nodes) return;
    const int n = t / 6, r = t - 6 * (t / 6);
    const int gr = 2 * n + r / 3, a = r % 3;
    const int r0 = gptr[gr], deg = gptr[gr + 1] - r0;
    const KT* row = Kb + 9 * (size_t)r0 + 3 * (size_t)deg * a;
    const bool has = (nflag[n] & 1) != 0;
    double out[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) out[k] = 0.0;
#pragma unroll
    for (int d = -1; d <= 1; ++d) {
        const int m = n + d;
        const bool present = d == 0 ? has : (d < 0 ? sv_link(nflag, m, n_nodes) : sv_link(nflag, n, n_nodes));
        if (!present) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int blk = sv_find(gptr, gcol, gr, 2 * m + h);
#pragma unroll
            for (int j = 0; j < 3; ++j) out[6 * (d + 1) + 3 * h + j] = (double)row[3 * (blk - r0) + j];
        }
    }
    const int di = 6 * n + r;
    const bool fi = fixed[di] != 0;
    double* o = Ab + 108 * (size_t)n + 18 * r;
#pragma unroll
    for (int k = 0; k < 18; ++k) {
        const int m = n - 1 + k / 6;
        const int dj = 6 * m + k % 6;
        double v = out[k];
        if (m >= 0 && m < n_nodes && (fi || fixed[dj])) v = (di == dj) ? 1.0 : 0.0;
        o[k] = v;
    }
}

// Block LU along each chain: S_n0 = D_n0, L_n = F_{n-1} S_{n-1}^-1,
// S_n = D_n - L_n E_{n-1}; stores L_n and S_n^-1 (Gauss-Jordan with partial
// pivoting, the first of the largest |entries| on ties).  One wavefront per
// chain: lane 6 i + j holds entry (i, j) of the node's 6x6 blocks, the block
// products and the elimination exchange entries by lane shuffles (no LDS, no
// barriers), and the next node's three entries per lane are loaded before
// the current node's elimination so that the loads overlap it.
__global__ __launch_bounds__(64) void sv_factor_kernel(const double* __restrict__ stop, const int* __restrict__ cstart,
                                                       const double* __restrict__ Ab, double* __restrict__ Lf,
                                                       double* __restrict__ Sinv, int* __restrict__ flags) {
    if (!nc_go(stop)) return;
    const int c = blockIdx.x;
    const int n0 = cstart[c], n1 = cstart[c + 1];
    const int l = threadIdx.x;
    const bool act = l < 36;
    const int i = act ? l / 6 : 0, j = act ? l - 6 * (l / 6) : 0;
    // node n's row i: [F (node n - 1) | D (node n) | E (node n + 1)], entry j of each
    auto load = [&](int n, double& D, double& F, double& E) {
        const double* a = Ab + 108 * (size_t)n + 18 * i;
        F = a[j];
        D = a[6 + j];
        E = a[12 + j];
    };
    double D, F, E, Dn = 0.0, Fn = 0.0, En = 0.0;
    load(n0, D, F, E);
    double P = 0.0, Ep = 0.0;   // S_{n-1}^-1 and E_{n-1}, entry (i, j)
    bool singular = false;
    for (int n = n0; n < n1; ++n) {
        if (n + 1 < n1) load(n + 1, Dn, Fn, En);
        double s = D;
        if (n > n0) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) t = fma(__shfl(F, 6 * i + k, 64), __shfl(P, 6 * k + j, 64), t);
            double u = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) u = fma(__shfl(t, 6 * i + k, 64), __shfl(Ep, 6 * k + j, 64), u);
            s -= u;
            if (act) Lf[36 * (size_t)n + l] = t;
        }
        // Gauss-Jordan on [S | I]: m holds S's entry, y the identity's
        double m = s, y = i == j ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            // pivot column entries come from fixed lanes: read lane to scalar
            int p = k;
            double best = fabs(sv_readlane(m, 7 * k));
#pragma unroll
            for (int r = k + 1; r < 6; ++r) {
                const double v = fabs(sv_readlane(m, 6 * r + k));
                if (v > best) {
                    best = v;
                    p = r;
                }
            }
            if (p != k) {   // uniform
                const double mk = __shfl(m, 6 * k + j, 64), mp = __shfl(m, 6 * p + j, 64);
                const double yk = __shfl(y, 6 * k + j, 64), yp = __shfl(y, 6 * p + j, 64);
                if (i == k) {
                    m = mp;
                    y = yp;
                } else if (i == p) {
                    m = mk;
                    y = yk;
                }
            }
            const double d = sv_readlane(m, 7 * k);
            if (d == 0.0) singular = true;
            if (i == k) {
                m = m / d;
                y = y / d;
            }
            const double f = __shfl(m, 6 * i + k, 64);
            const double pm = __shfl(m, 6 * k + j, 64), py = __shfl(y, 6 * k + j, 64);
            if (i != k) {
                m -= f * pm;
                y -= f * py;
            }
        }
        if (act) Sinv[36 * (size_t)n + l] = y;
        P = y;
        Ep = E;
        D = Dn;
        F = Fn;
        E = En;
    }
    if (singular && l == 0) atomicOr(flags, XK_FLAG_SINGULAR);
}

__global__ void sv_slist_kernel(const double* __restrict__ stop,
                                int64_t n_pairs, const int* __restrict__ sinc, int* __restrict__ slist) {
    if (!nc_go(stop)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pairs && sinc[i] >= 0) slist[sinc[i]] = (int)i;
}

__global__ void sv_slot_nodes_kernel(const double* __restrict__ stop,
                                     int ns, const int* __restrict__ slist, PairSoA P, int4* __restrict__ sn) {
    if (!nc_go(stop)) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ns) return;
    const int i = slist[q];
    sn[q] = make_int4(P.na0[i], P.na1[i], P.nb0[i], P.nb1[i]);
}

// the pair's staged tangent record -> G (3x3) and c (4) in FP64
template <class GT>
__global__ void sv_slot_pack_kernel(const double* __restrict__ stop,
                                    int ns, const GT* __restrict__ sg, double* __restrict__ sG,
                                    double* __restrict__ sc) {
    if (!nc_go(stop)) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ns) return;
    const GT* r = sg + SG_REC * (size_t)q;
#pragma unroll
    for (int k = 0; k < 9; ++k) sG[9 * (size_t)q + k] = (double)r[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) sc[4 * (size_t)q + k] = (double)r[SG_C + k];
}

// Compliance of each chain at its contact nodes' translational dofs:
// C_w[(k, a), (l, b)] = (A~_b^-1)[(node k, a), (node l, b)].  One workgroup per
// chain, one lane per right-hand side e_(l, b); forward sweep into Y, backward
// sweep reading it.
__global__ __launch_bounds__(256) void sv_compliance_kernel(
    const double* __restrict__ stop, const int* __restrict__ cstart, const int* __restrict__ cn,
    const int* __restrict__ cnode, const int* __restrict__ ch_cn0, const int* __restrict__ ch_m,
    const int64_t* __restrict__ ch_yoff, const int64_t* __restrict__ ch_coff, const uint8_t* __restrict__ fixed,
    const double* __restrict__ Ab, const double* __restrict__ Lf, const double* __restrict__ Sinv,
    double* __restrict__ Y, double* __restrict__ C) {
    if (!nc_go(stop)) return;
    const int c = blockIdx.x;
    const int m = ch_m[c];
    if (m == 0) return;
    const int n0 = cstart[c], n1 = cstart[c + 1], cn0 = ch_cn0[c];
    double* Yc = Y + ch_yoff[c];
    double* Cc = C + ch_coff[c];
    for (int rho = threadIdx.x; rho < m; rho += blockDim.x) {
        const int s = cnode[cn0 + rho / 3], comp = rho % 3;
        const bool live = fixed[6 * s + comp] == 0;
        double y[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) y[a] = (live && a == comp) ? 1.0 : 0.0;
        for (int n = s; n < n1; ++n) {
            if (n > s) {
                const double* L = Lf + 36 * (size_t)n;
                double t[6];
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    double v = 0.0;
#pragma unroll
                    for (int b = 0; b < 6; ++b) v = fma(L[6 * a + b], y[b], v);
                    t[a] = -v;
                }
#pragma unroll
                for (int a = 0; a < 6; ++a) y[a] = t[a];
            }
#pragma unroll
            for (int a = 0; a < 6; ++a) Yc[((size_t)(n - n0) * 6 + a) * m + rho] = y[a];
        }
        double x[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) x[a] = 0.0;
        for (int n = n1 - 1; n >= n0; --n) {
            double v[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) v[a] = n >= s ? Yc[((size_t)(n - n0) * 6 + a) * m + rho] : 0.0;
            if (n < n1 - 1) {
                const double* e = Ab + 108 * (size_t)n + 12;
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    double t = 0.0;
#pragma unroll
                    for (int b = 0; b < 6; ++b) t = fma(e[18 * a + b], x[b], t);
                    v[a] -= t;
                }
            }
            const double* Si = Sinv + 36 * (size_t)n;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < 6; ++b) t = fma(Si[6 * a + b], v[b], t);
                x[a] = t;
            }
            const int k = cn[n];
            if (k >= 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) Cc[(size_t)(3 * (k - cn0) + a) * m + rho] = x[a];
            }
        }
    }
}

// Capacitance block (i, q) of one component, column-major with ld = 3 Pc:
// I delta_iq + G_i sum_{k,l} c_ik c_ql M_k C[node_k(i), node_l(q)] M_l.
__global__ void sv_capacitance_kernel(const double* __restrict__ stop,
                                      int Pc, const int* __restrict__ cslot, const int4* __restrict__ sn,
                                      const double* __restrict__ sG, const double* __restrict__ sc,
                                      const int* __restrict__ node_chain, const int* __restrict__ cn,
                                      const int* __restrict__ ch_cn0, const int* __restrict__ ch_m,
                                      const int64_t* __restrict__ ch_coff, const uint8_t* __restrict__ fixed,
                                      const double* __restrict__ C, double* __restrict__ S) {
    if (!nc_go(stop)) return;
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const int lq = blockIdx.y;
    if (li >= Pc) return;
    const int i = cslot[li], q = cslot[lq];
    const int4 a4 = sn[i], b4 = sn[q];
    const int na[4] = {a4.x, a4.y, a4.z, a4.w};
    const int nb[4] = {b4.x, b4.y, b4.z, b4.w};
    double B[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) B[k] = 0.0;
    for (int k = 0; k < 4; ++k) {
        const int a = na[k], ch = node_chain[a];
        const double ck = sc[4 * (size_t)i + k];
        for (int l = 0; l < 4; ++l) {
            const int b = nb[l];
            if (node_chain[b] != ch) continue;
            const int m = ch_m[ch], cn0 = ch_cn0[ch];
            const double* Cb = C + ch_coff[ch] + (size_t)(3 * (cn[a] - cn0)) * m + 3 * (cn[b] - cn0);
            const double w = ck * sc[4 * (size_t)q + l];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (fixed[6 * a + r]) continue;
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    if (!fixed[6 * b + s]) B[3 * r + s] = fma(w, Cb[(size_t)r * m + s], B[3 * r + s]);
            }
        }
    }
    const double* G = sG + 9 * (size_t)i;
    const size_t ld = 3 * (size_t)Pc;
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double v = (li == lq && r == s) ? 1.0 : 0.0;
#pragma unroll
            for (int t = 0; t < 3; ++t) v = fma(G[3 * r + t], B[3 * t + s], v);
            S[(3 * (size_t)lq + s) * ld + 3 * (size_t)li + r] = v;
        }
}

__global__ void sv_info_kernel(const double* __restrict__ stop,
                               int nc, const int64_t* __restrict__ info, int* __restrict__ flags) {
    if (!nc_go(stop)) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc && info[c] != 0) atomicOr(flags, XK_FLAG_SINGULAR);
}

// ------------------------------------------------------- capacitance LU
// Blocked right-looking LU without row interchanges of one column-major
// n x n capacitance block (leading dimension n), panels of LU_NB columns.
// Per panel one launch: every workgroup factors the w x w diagonal block in
// LDS (the same arithmetic in each, so the copies agree bitwise; workgroup 0
// stores it), then forms either LU_NB rows of L21 = A21 U11^-1 or LU_NB
// columns of U12 = L11^-1 A12 in an LDS tile; rocBLAS's DGEMM then applies
// A22 -= L21 U12 on the matrix cores.  Inside the launch the diagonal block
// and the tiles are processed in sub-blocks of LU_SB: the sub-block's
// factorization or triangular solve runs with one row (or column) per lane in
// registers, the rest of the block is updated by 3 x 3 / 4 x 3 register tiles
// (the code stays a few thousand instructions: the launch is not bound by
// instruction fetch; full panels use a variant whose unrolled steps carry no
// bound checks).  Divisions by a pivot are products with its reciprocal
// (formed once per pivot), which keeps the divisions off the dependent chains.
//
// Without interchanges the pivots are the Schur complements' diagonals in
// order.  On the strands' capacitance blocks (far from diagonally dominant,
// condition numbers near 1e12) this LU had growth max|U| / max|C| = 25
// against 1.7 for partial pivoting, and one refinement step brought both to
// the same residual (measured at 91 wires); the solve refines against the
// assembled operator in any case.  A pivot that is zero or not finite is
// reported in info (first such column, 1-based), and the caller then
// factors the block again with partial pivoting.
constexpr int LU_NB = 64;   // panel width = tile size
constexpr int LU_SB = 16;   // sub-block inside the panel
constexpr int LU_NP = 2;    // panels per outer panel (one trailing update each)


// FULL: the panel holds LU_NB columns (every panel but a last partial one):
// the sub-block bounds are constants and the unrolled steps carry no guards.
template <bool FULL>
__global__ __launch_bounds__(256) void sv_lu_panel_kernel(const double* __restrict__ stop,
                                                          int n, int j0, int n_row_wg, double* __restrict__ A,
                                                          int64_t* __restrict__ info) {
    if (!nc_go(stop)) return;
    __shared__ double D[LU_NB][LU_NB + 1];
    __shared__ double X[LU_NB][LU_NB + 1];
    __shared__ double invd[LU_NB];   // reciprocals of the pivots
    __shared__ int s_bad;
    const int w = FULL ? LU_NB : min(LU_NB, n - j0);
    const int tid = threadIdx.x;
    const size_t ld = (size_t)n;
    constexpr int CP = LU_NB * LU_NB / 256;   // tile entries per lane (all loads issued before the stores)
#pragma unroll
    for (int it = 0; it < CP; ++it) {
        const int k = tid + 256 * it, r = k & (LU_NB - 1), q = k / LU_NB;
        D[r][q] = (r < w && q < w) ? A[(size_t)(j0 + r) + (size_t)(j0 + q) * ld] : 0.0;
    }
    // the workgroup's tile (rows below or columns right of the panel, which the
    // diagonal block's factorization does not touch) is loaded now, so that its
    // loads overlap the diagonal block's
    const int rest = n - j0 - w;
    const bool row_tile = (int)blockIdx.x < n_row_wg;
    const int row0 = j0 + w + (int)blockIdx.x * LU_NB, mr = min(LU_NB, n - row0);
    const int col0 = j0 + w + ((int)blockIdx.x - n_row_wg) * LU_NB, nc = min(LU_NB, n - col0);
    if (rest > 0) {   // uniform
        if (row_tile) {
            // X[i][k] = A[row0 + i, j0 + k]
#pragma unroll
            for (int it = 0; it < CP; ++it) {
                const int k = tid + 256 * it, i = k & (LU_NB - 1), q = k / LU_NB;
                X[i][q] = (i < mr && q < w) ? A[(size_t)(row0 + i) + (size_t)(j0 + q) * ld] : 0.0;
            }
        } else {
            // X[k][j] = A[j0 + k, col0 + j]
#pragma unroll
            for (int it = 0; it < CP; ++it) {
                const int k = tid + 256 * it, q = k & (LU_NB - 1), j = k / LU_NB;
                X[q][j] = (q < w && j < nc) ? A[(size_t)(j0 + q) + (size_t)(col0 + j) * ld] : 0.0;
            }
        }
    }
    if (tid == 0) s_bad = -1;
    __syncthreads();
    // ---- diagonal block
#pragma unroll 1
    for (int o = 0; o < w; o += LU_SB) {
        const int bw = FULL ? LU_SB : min(LU_SB, w - o);
        if (tid < 64) {
            // (a) the bw x bw sub-block: lane i holds row o + i, the pivot row
            // of step c broadcast from lane c
            const int i = tid;
            double v[LU_SB];
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) v[k] = (i < bw && k < bw) ? D[o + i][o + k] : 0.0;
            // lane c keeps pivot c's reciprocal and validity in registers (no LDS
            // traffic and no branch on the chain from one pivot to the next)
            double my_ip = 0.0;
            bool my_bad = false;
#pragma unroll
            for (int c = 0; c < LU_SB; ++c) {
                if (c < bw) {   // uniform
                    const double p = sv_readlane(v[c], c);
                    const double ip = 1.0 / p;
                    const bool bad = !(fabs(p) > 0.0 && fabs(p) <= DBL_MAX);
                    my_ip = i == c ? ip : my_ip;
                    my_bad = i == c ? bad : my_bad;
                    const bool below = i > c && i < bw;
                    const double l = below ? v[c] * ip : 0.0;   // rows at or above c: unchanged
#pragma unroll
                    for (int k = c + 1; k < LU_SB; ++k) v[k] = fma(-l, sv_readlane(v[k], c), v[k]);
                    if (below) v[c] = l;
                }
            }
            if (i < bw) invd[o + i] = my_ip;
            const unsigned long long badm = __ballot(i < bw && my_bad);
            if (i == 0 && badm != 0ull && s_bad < 0) s_bad = o + __ffsll((long long)badm) - 1;   // the first one
            if (i < bw)
#pragma unroll
                for (int k = 0; k < LU_SB; ++k)
                    if (k < bw) D[o + i][o + k] = v[k];
        }
        __syncthreads();
        const int R = w - o - LU_SB;   // rows / columns after the sub-block
        if (R <= 0) break;             // uniform
        if (tid < R) {
            // (b) L part of the rows below: x U_kk^-1
            const int r = o + LU_SB + tid;
            double x[LU_SB];
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) x[k] = D[r][o + k];
#pragma unroll
            for (int c = 0; c < LU_SB; ++c) {
                x[c] = x[c] * invd[o + c];
#pragma unroll
                for (int k = c + 1; k < LU_SB; ++k) x[k] = fma(-x[c], D[o + c][o + k], x[k]);
            }
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) D[r][o + k] = x[k];
        } else if (tid >= 64 && tid < 64 + R) {
            // (b) U part of the columns right of it: L_kk^-1 y
            const int q = o + LU_SB + (tid - 64);
            double y[LU_SB];
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) y[k] = D[o + k][q];
#pragma unroll
            for (int c = 0; c < LU_SB; ++c)
#pragma unroll
                for (int k = c + 1; k < LU_SB; ++k) y[k] = fma(-D[o + k][o + c], y[c], y[k]);
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) D[o + k][q] = y[k];
        }
        __syncthreads();
        {
            // (c) trailing update of the R x R rest: 3 x 3 tiles (R <= 48)
            const int bi = tid / 16, bj = tid % 16;
            const int r0 = o + LU_SB + 3 * bi, q0 = o + LU_SB + 3 * bj;
            if (3 * bi < R && 3 * bj < R) {
                // rows / columns past R (a partial tile) read as zero, are not stored
                bool ra[3], cb[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    ra[a] = 3 * bi + a < R;
                    cb[a] = 3 * bj + a < R;
                }
                double acc[3][3];
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) acc[a][b] = (ra[a] && cb[b]) ? D[r0 + a][q0 + b] : 0.0;
#pragma unroll
                for (int k = 0; k < LU_SB; ++k) {
                    double lk[3], uk[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) lk[a] = ra[a] ? D[r0 + a][o + k] : 0.0;
#pragma unroll
                    for (int b = 0; b < 3; ++b) uk[b] = cb[b] ? D[o + k][q0 + b] : 0.0;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) acc[a][b] = fma(-lk[a], uk[b], acc[a][b]);
                }
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b)
                        if (ra[a] && cb[b]) D[r0 + a][q0 + b] = acc[a][b];
            }
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) {
#pragma unroll
        for (int it = 0; it < CP; ++it) {
            const int k = tid + 256 * it, r = k & (LU_NB - 1), q = k / LU_NB;
            if (r < w && q < w) A[(size_t)(j0 + r) + (size_t)(j0 + q) * ld] = D[r][q];
        }
        if (tid == 0) {
            int64_t v = j0 == 0 ? 0 : *info;
            if (v == 0 && s_bad >= 0) v = j0 + s_bad + 1;
            *info = v;
        }
    }
    if (rest <= 0) return;
    if (row_tile) {
        // ---- LU_NB rows of L21: X U11^-1, X[i][k] = A[row0 + i, j0 + k]
#pragma unroll 1
        for (int o = 0; o < w; o += LU_SB) {
            const int bw = FULL ? LU_SB : min(LU_SB, w - o);
            if (tid < LU_NB) {
                double x[LU_SB];
#pragma unroll
                for (int k = 0; k < LU_SB; ++k) x[k] = X[tid][o + k];
#pragma unroll
                for (int c = 0; c < LU_SB; ++c) {
                    if (c < bw) {
                        x[c] = x[c] * invd[o + c];
#pragma unroll
                        for (int k = c + 1; k < LU_SB; ++k)
                            if (k < bw) x[k] = fma(-x[c], D[o + c][o + k], x[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < LU_SB; ++k) X[tid][o + k] = x[k];
            }
            __syncthreads();
            const int R = w - o - LU_SB;
            if (R <= 0) break;
            {
                // X[:, o + 16 ..] -= X[:, o .. o + 16) U[o .. o + 16), o + 16 ..]: 4 x 3 tiles
                const int bi = tid / 16, bj = tid % 16;
                const int i0 = 4 * bi, q0 = o + LU_SB + 3 * bj;
                if (3 * bj < R) {
                    bool cb[3];
#pragma unroll
                    for (int b = 0; b < 3; ++b) cb[b] = 3 * bj + b < R;
                    double acc[4][3];
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) acc[a][b] = cb[b] ? X[i0 + a][q0 + b] : 0.0;
#pragma unroll
                    for (int k = 0; k < LU_SB; ++k) {
                        double xk[4], uk[3];
#pragma unroll
                        for (int a = 0; a < 4; ++a) xk[a] = X[i0 + a][o + k];
#pragma unroll
                        for (int b = 0; b < 3; ++b) uk[b] = cb[b] ? D[o + k][q0 + b] : 0.0;
#pragma unroll
                        for (int a = 0; a < 4; ++a)
#pragma unroll
                            for (int b = 0; b < 3; ++b) acc[a][b] = fma(-xk[a], uk[b], acc[a][b]);
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b)
                            if (cb[b]) X[i0 + a][q0 + b] = acc[a][b];
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int it = 0; it < CP; ++it) {
            const int k = tid + 256 * it, i = k & (LU_NB - 1), q = k / LU_NB;
            if (i < mr && q < w) A[(size_t)(row0 + i) + (size_t)(j0 + q) * ld] = X[i][q];
        }
        return;
    }
    // ---- LU_NB columns of U12: L11^-1 Y, X[k][j] = A[j0 + k, col0 + j]
#pragma unroll 1
    for (int o = 0; o < w; o += LU_SB) {
        const int bw = FULL ? LU_SB : min(LU_SB, w - o);
        if (tid < LU_NB) {
            double y[LU_SB];
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) y[k] = X[o + k][tid];
#pragma unroll
            for (int c = 0; c < LU_SB; ++c) {
                if (c < bw) {
#pragma unroll
                    for (int k = c + 1; k < LU_SB; ++k)
                        if (k < bw) y[k] = fma(-D[o + k][o + c], y[c], y[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < LU_SB; ++k) X[o + k][tid] = y[k];
        }
        __syncthreads();
        const int R = w - o - LU_SB;
        if (R <= 0) break;
        {
            // X[o + 16 .., :] -= L[o + 16 .., o .. o + 16) X[o .. o + 16), :]: 3 x 4 tiles
            const int bi = tid % 16, bj = tid / 16;
            const int r0 = o + LU_SB + 3 * bi, j0t = 4 * bj;
            if (3 * bi < R) {
                bool ra[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) ra[a] = 3 * bi + a < R;
                double acc[3][4];
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = ra[a] ? X[r0 + a][j0t + b] : 0.0;
#pragma unroll
                for (int k = 0; k < LU_SB; ++k) {
                    double lk[3], yk[4];
#pragma unroll
                    for (int a = 0; a < 3; ++a) lk[a] = ra[a] ? D[r0 + a][o + k] : 0.0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) yk[b] = X[o + k][j0t + b];
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) acc[a][b] = fma(-lk[a], yk[b], acc[a][b]);
                }
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (ra[a]) X[r0 + a][j0t + b] = acc[a][b];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < CP; ++it) {
        const int k = tid + 256 * it, q = k & (LU_NB - 1), j = k / LU_NB;
        if (q < w && j < nc) A[(size_t)(j0 + q) + (size_t)(col0 + j) * ld] = X[q][j];
    }
}

// pivots of an LU without interchanges: row i stays row i (LAPACK, 1-based)
__global__ void sv_ipiv_identity_kernel(const double* __restrict__ stop, int64_t n, int64_t* __restrict__ ipiv) {
    if (!nc_go(stop)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ipiv[i] = i + 1;
}

constexpr int CS_PF = 4;   // nodes in flight in the chain sweeps

// x <- A~_b^-1 b along every chain (b and x may be the same vector): 8 lanes
// per chain (lanes 0..5 hold the node's six components), 8 chains per
// wavefront; rows of the 6x6 blocks are combined across the chain's lanes by
// shuffles.
__global__ __launch_bounds__(64) void sv_chain_solve_kernel(const double* __restrict__ stop, int n_chains,
                                                            const int* __restrict__ cstart,
                                                            const double* __restrict__ Ab,
                                                            const double* __restrict__ Lf,
                                                            const double* __restrict__ Sinv, const double* b,
                                                            double* x) {
    if (!nc_go(stop)) return;
    const int lane = threadIdx.x, base = lane & ~7, a = lane & 7;
    const int c = blockIdx.x * 8 + (lane >> 3);
    if (c >= n_chains || a >= 6) return;
    const int n0 = cstart[c], n1 = cstart[c + 1];
    // forward: y_n = x_n - L_n y_{n-1}.  The rows and values of the next
    // CS_PF nodes are in flight while node n's shuffles run (a ring of
    // registers, slot q holding node n0 + q mod CS_PF): one node ahead left
    // each step waiting on its loads.
    double y = 0.0;
    double xr[CS_PF], Lr[CS_PF][6];
#pragma unroll
    for (int q = 0; q < CS_PF; ++q) {
        const int m = n0 + q;
        xr[q] = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) Lr[q][b] = 0.0;
        if (m < n1) {
            xr[q] = b[6 * (size_t)m + a];
            if (m > n0) {
                const double* p = Lf + 36 * (size_t)m + 6 * a;
#pragma unroll
                for (int b = 0; b < 6; ++b) Lr[q][b] = p[b];
            }
        }
    }
    for (int n = n0; n < n1; n += CS_PF) {
#pragma unroll
        for (int q = 0; q < CS_PF; ++q) {
            const int m = n + q;
            if (m < n1) {
                double v = xr[q];
                if (m > n0) {
                    double t = 0.0;
#pragma unroll
                    for (int b = 0; b < 6; ++b) t = fma(Lr[q][b], __shfl(y, base + b, 64), t);
                    v -= t;
                }
                y = v;
                x[6 * (size_t)m + a] = v;
                const int mf = m + CS_PF;   // refill the slot
                if (mf < n1) {
                    xr[q] = b[6 * (size_t)mf + a];
                    const double* p = Lf + 36 * (size_t)mf + 6 * a;
#pragma unroll
                    for (int b = 0; b < 6; ++b) Lr[q][b] = p[b];
                }
            }
        }
    }
    // backward: x_n = S_n^-1 (y_n - E_n x_{n+1}), the same ring from the last
    // node down (slot q holding node n1 - 1 - q mod CS_PF)
    double z = 0.0;
    double yr[CS_PF], Er[CS_PF][6], Sr[CS_PF][6];
#pragma unroll
    for (int q = 0; q < CS_PF; ++q) {
        const int m = n1 - 1 - q;
        yr[q] = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) Er[q][b] = Sr[q][b] = 0.0;
        if (m >= n0) {
            yr[q] = x[6 * (size_t)m + a];
            const double* pe = Ab + 108 * (size_t)m + 18 * a + 12;
            const double* ps = Sinv + 36 * (size_t)m + 6 * a;
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                Er[q][b] = pe[b];
                Sr[q][b] = ps[b];
            }
        }
    }
    for (int n = n1 - 1; n >= n0; n -= CS_PF) {
#pragma unroll
        for (int q = 0; q < CS_PF; ++q) {
            const int m = n - q;
            if (m >= n0) {
                double v = yr[q];
                if (m < n1 - 1) {
                    double t = 0.0;
#pragma unroll
                    for (int b = 0; b < 6; ++b) t = fma(Er[q][b], __shfl(z, base + b, 64), t);
                    v -= t;
                }
                double w = 0.0;
#pragma unroll
                for (int b = 0; b < 6; ++b) w = fma(Sr[q][b], __shfl(v, base + b, 64), w);
                z = w;
                x[6 * (size_t)m + a] = w;
                const int mf = m - CS_PF;
                if (mf >= n0) {
                    yr[q] = x[6 * (size_t)mf + a];
                    const double* pe = Ab + 108 * (size_t)mf + 18 * a + 12;
                    const double* ps = Sinv + 36 * (size_t)mf + 6 * a;
#pragma unroll
                    for (int b = 0; b < 6; ++b) {
                        Er[q][b] = pe[b];
                        Sr[q][b] = ps[b];
                    }
                }
            }
        }
    }
}

// ---- dense LU solves of the capacitance blocks (column-major, ld = n).
// Substitution, right-looking, in row blocks of TB: step s updates the rows
// below block s with block s's solved values (workgroup 0 takes block s + 1
// and then solves its diagonal block with the block's rows held in
// registers).  Backward stable, unlike solves through inverted diagonal blocks.
constexpr int TB = 64;

// acc - sum_{k < m} A[i + (k0 + k) ld] ys[k], all loads in flight for full blocks
__device__ __forceinline__ double sv_block_dot(double acc, const double* __restrict__ A, size_t ld, int i, int k0,
                                               int m, const double* ys) {
    const double* a = A + i + (size_t)k0 * ld;
    if (m == TB) {
        double p[TB];
#pragma unroll
        for (int k = 0; k < TB; ++k) p[k] = a[(size_t)k * ld];
#pragma unroll
        for (int k = 0; k < TB; ++k) acc = fma(-p[k], ys[k], acc);
    } else {
        for (int k = 0; k < m; ++k) acc = fma(-a[(size_t)k * ld], ys[k], acc);
    }
    return acc;
}

// The row interchanges of one capacitance block (LAPACK pivots) as a
// permutation, perm[i] = original row that ends in row i: one workgroup, the
// sequential interchanges in LDS (blocks up to PERM_MAX rows), so the solve
// needs no host round trip (and can be captured in a graph).
constexpr int PERM_MAX = 16384;

__global__ __launch_bounds__(256) void sv_perm_kernel(const double* __restrict__ stop,
                                                      int n, const int64_t* __restrict__ ipiv, int* __restrict__ perm,
                                                      int* __restrict__ flags) {
    if (!nc_go(stop)) return;
    __shared__ int p[PERM_MAX];
    __shared__ int q[PERM_MAX];
    bool bad = false;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        p[i] = i;
        const int64_t v = ipiv[i] - 1;
        bad |= v < i || v >= n;   // partial pivoting picks a row at or below the diagonal
        q[i] = (int)v;
    }
    __syncthreads();
    if (bad) atomicOr(flags, XK_FLAG_SINGULAR);
    if (threadIdx.x == 0) {
        for (int k = 0; k < n; ++k) {
            const int r = q[k];
            if (r < k || r >= n) continue;
            const int a = p[k];
            p[k] = p[r];
            p[r] = a;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = p[i];
}

// y[i] = t[perm[i]]
__global__ void sv_permute_kernel(const double* __restrict__ stop,
                                  int n, const int* __restrict__ perm, const double* __restrict__ t,
                                  double* __restrict__ y) {
    if (!nc_go(stop)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = t[perm[i]];
}


// The triangular solves advance TS blocks per launch.  Unit lower triangle,
// steps s0 .. s0 + ks - 1: every workgroup (one wave) solves those diagonal
// blocks of y in order -- each block's rows first take the updates of the
// blocks solved before it in this launch, then forward substitution with the
// solved values broadcast from lane to lane -- keeping the solutions in LDS
// (workgroup 0 stores them in z); every workgroup then subtracts
// L[i, block s0 + j] z_{s0 + j}, j = 0 .. ks - 1 in order, from its rows i of
// block s0 + ks + blockIdx.x.  Every row sees the updates in the order of the
// one-block-per-launch form, so the result does not depend on TS; the
// redundant diagonal work of the workgroups reads the same blocks from L2,
// and fewer launches are the saving: up to TS blocks per launch, TS for small
// blocks (the 91-wire squeeze, 21 blocks: 2.12 -> 2.00 ms per attempt), half of
// it for large ones, where the redundant diagonal work of many workgroups
// costs about what the launches save (6,012 rows, 94 blocks: 15.1 -> 14.8 ms
// per solve with 2, 15.2 with 4).
constexpr int TS = 4;
static int sv_tri_steps(int nb) { return nb <= 32 ? TS : TS / 2; }

__global__ __launch_bounds__(TB) void sv_lower_step_kernel(const double* __restrict__ stop, int n,
                                                           const double* __restrict__ LU,
                                                           double* __restrict__ y, double* __restrict__ z,
                                                           int s0, int ks) {
    if (!nc_go(stop)) return;
    __shared__ double zs[TS][TB];
    const int t = threadIdx.x;
    const size_t ld = (size_t)n;
    for (int j = 0; j < ks; ++j) {   // ks <= TS, uniform
        const int k0 = (s0 + j) * TB, m = min(TB, n - k0);
        double v = t < m ? y[k0 + t] : 0.0;
        for (int i = 0; i < j; ++i) v = sv_block_dot(v, LU, ld, t < m ? k0 + t : k0, (s0 + i) * TB, TB, zs[i]);
        if (t >= m) v = 0.0;
        double row[TB];
#pragma unroll
        for (int k = 0; k < TB; ++k) row[k] = (t < m && k < t) ? LU[(size_t)(k0 + t) + (size_t)(k0 + k) * ld] : 0.0;
#pragma unroll
        for (int k = 0; k < TB; ++k) v = fma(-row[k], sv_readlane(v, k), v);   // row[k] = 0 for k >= t
        zs[j][t] = v;
        if (blockIdx.x == 0 && t < m) z[k0 + t] = v;
        __syncthreads();
    }
    const int i = (s0 + ks + blockIdx.x) * TB + t;
    if (i >= n) return;
    double acc = y[i];
    for (int j = 0; j < ks; ++j)
        acc = sv_block_dot(acc, LU, ld, i, (s0 + j) * TB, min(TB, n - (s0 + j) * TB), zs[j]);
    y[i] = acc;
}

// Upper triangle with its diagonal, steps s0, s0 - 1, .., s0 - ks + 1 (from
// the last block down): every workgroup solves those diagonal blocks of z in
// order by back substitution (each after the updates of the blocks solved
// before it in this launch; workgroup 0 stores them in x), then subtracts
// U[i, block s0 - j] x_{s0 - j}, j = 0 .. ks - 1 in order, from its rows i of
// block s0 - ks - blockIdx.x.
__global__ __launch_bounds__(TB) void sv_upper_step_kernel(const double* __restrict__ stop, int n,
                                                           const double* __restrict__ LU,
                                                           double* __restrict__ z, double* __restrict__ x,
                                                           int s0, int ks) {
    if (!nc_go(stop)) return;
    __shared__ double xs[TS][TB];
    const int t = threadIdx.x;
    const size_t ld = (size_t)n;
    for (int j = 0; j < ks; ++j) {
        const int k0 = (s0 - j) * TB, m = min(TB, n - k0);
        double v = t < m ? z[k0 + t] : 0.0;
        for (int i = 0; i < j; ++i)
            v = sv_block_dot(v, LU, ld, t < m ? k0 + t : k0, (s0 - i) * TB, min(TB, n - (s0 - i) * TB), xs[i]);
        // back substitution on the row scaled by the reciprocal of its
        // diagonal: x_t = z_t / u_tt - sum_{k > t} (u_tk / u_tt) x_k, so step k
        // reads lane k's finished value with no multiply in the chain
        const double idg = 1.0 / (t < m ? LU[(size_t)(k0 + t) + (size_t)(k0 + t) * ld] : 1.0);
        v = t < m ? v * idg : 0.0;
        double row[TB];
#pragma unroll
        for (int k = 0; k < TB; ++k)
            row[k] = (t < m && k > t && k < m) ? LU[(size_t)(k0 + t) + (size_t)(k0 + k) * ld] : 0.0;
        // (the scaling of row[k] is off the chain through v, and after the loads:
        // scaling where the masked loads are issued would wait on each load)
#pragma unroll
        for (int k = TB - 1; k >= 0; --k) v = fma(-(row[k] * idg), sv_readlane(v, k), v);   // row[k] = 0 for k <= t
        xs[j][t] = v;
        if (blockIdx.x == 0 && t < m) x[k0 + t] = v;
        __syncthreads();
    }
    const int b = s0 - ks - (int)blockIdx.x;
    if (b < 0) return;
    const int i = b * TB + t;
    double acc = z[i];
    for (int j = 0; j < ks; ++j)
        acc = sv_block_dot(acc, LU, ld, i, (s0 - j) * TB, min(TB, n - (s0 - j) * TB), xs[j]);
    z[i] = acc;
}

__global__ void sv_perm_identity_kernel(const double* __restrict__ stop, int n, int* __restrict__ perm) {
    if (!nc_go(stop)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[i] = i;
}

// W[3 pos(q) + r] = (G_q U~_q^T y)_r
__global__ void sv_project_kernel(const double* __restrict__ stop,
                                  int ns, const int* __restrict__ pos, const int4* __restrict__ sn,
                                  const double* __restrict__ sG, const double* __restrict__ sc,
                                  const uint8_t* __restrict__ fixed, const double* __restrict__ y,
                                  double* __restrict__ W) {
    if (!nc_go(stop)) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ns) return;
    const int4 n4 = sn[q];
    const int nd[4] = {n4.x, n4.y, n4.z, n4.w};
    double u[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double ck = sc[4 * (size_t)q + k];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const size_t d = 6 * (size_t)nd[k] + a;
            if (!fixed[d]) u[a] = fma(ck, y[d], u[a]);
        }
    }
    const double* G = sG + 9 * (size_t)q;
#pragma unroll
    for (int r = 0; r < 3; ++r) W[3 * (size_t)pos[q] + r] = G[3 * r] * u[0] + G[3 * r + 1] * u[1] + G[3 * r + 2] * u[2];
}

// out = r - U~ W  (node lists hold the codes 4 slot + k of the pairs at each node)
__global__ void sv_lift_kernel(const double* __restrict__ stop,
                               int n_nodes, const int* __restrict__ pptr, const int* __restrict__ psrc,
                               const int* __restrict__ pos, const double* __restrict__ sc,
                               const uint8_t* __restrict__ fixed, const double* __restrict__ r,
                               const double* __restrict__ W, double* __restrict__ out) {
    if (!nc_go(stop)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * n_nodes) return;
    const int n = t / 6, a = t - 6 * (t / 6);
    double v = r[t];
    if (a < 3 && !fixed[t]) {
        for (int p = pptr[n]; p < pptr[n + 1]; ++p) {
            const int code = psrc[p], q = code >> 2, k = code & 3;
            v = fma(-sc[4 * (size_t)q + k], W[3 * (size_t)pos[q] + a], v);
        }
    }
    out[t] = v;
}

// out = b - A~_b x from the chain blocks (one lane per scalar row); the
// contact part U~ G U~^T x is subtracted by sv_project_kernel + sv_lift_kernel
__global__ void sv_chain_residual_kernel(const double* __restrict__ stop,
                                         int n_nodes, const double* __restrict__ Ab, const double* __restrict__ b,
                                         const double* __restrict__ x, double* __restrict__ out) {
    if (!nc_go(stop)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * n_nodes) return;
    const int n = t / 6, r = t - 6 * (t / 6);
    const double* a = Ab + 108 * (size_t)n + 18 * r;
    double v = 0.0;
    const int k0 = n > 0 ? 0 : 6, k1 = n + 1 < n_nodes ? 18 : 12;
    for (int k = k0; k < k1; ++k) v = fma(a[k], x[6 * (size_t)(n - 1) + k], v);
    out[t] = b[t] - v;
}

__global__ void sv_rhs_kernel(const double* __restrict__ stop,
                              int ndof, const double* __restrict__ R, const uint8_t* __restrict__ fixed,
                              double* __restrict__ b) {
    if (!nc_go(stop)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ndof) b[i] = fixed[i] ? 0.0 : -R[i];
}

__global__ void sv_axpy_kernel(const double* __restrict__ stop,
                               int ndof, const double* __restrict__ dx, double* __restrict__ x) {
    if (!nc_go(stop)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ndof) x[i] += dx[i];
}

// rel = ||r|| / ||b|| (one workgroup; fixed order, so deterministic)
__global__ __launch_bounds__(1024) void sv_relnorm_kernel(const double* __restrict__ stop,
                                                          int ndof, const double* __restrict__ r,
                                                          const double* __restrict__ b, double* __restrict__ rel) {
    if (!nc_go(stop)) return;
    __shared__ double s0[1024], s1[1024];
    double a = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < ndof; i += 1024) {
        a = fma(r[i], r[i], a);
        c = fma(b[i], b[i], c);
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *rel = s1[0] > 0.0 ? sqrt(s0[0] / s1[0]) : sqrt(s0[0]);
}

// ------------------------------------------------------------------ host

// Device-decided refinement (a control record attached, tol > 0): the gate
// word XK_NC_RGATE of the record is 0 while steps remain to be taken; the
// test after each step's residual norm closes it at the tolerance, at the
// rounding floor (a step no longer halves the residual) or on a non-finite
// norm -- the host loop's rule -- and the closed gate turns the remaining
// steps' kernels into no-ops.  A call that has ended never opens it.
__global__ void sv_refine_begin_kernel(double* __restrict__ c) {
    if (threadIdx.x != 0) return;
    c[XK_NC_RGATE] = c[XK_NC_STATE] == XK_NC_RUNNING ? 0.0 : 1.0;
    c[XK_NC_RPREV] = HUGE_VAL;
    c[XK_NC_RSTEPS] = 0.0;
}

__global__ void sv_refine_test_kernel(double* __restrict__ c, const double* __restrict__ red, double tol) {
    if (threadIdx.x != 0 || c[XK_NC_RGATE] != 0.0) return;
    const double rel = *red;
    if (rel <= tol || rel > 0.5 * c[XK_NC_RPREV] || !(rel <= DBL_MAX)) {
        c[XK_NC_RGATE] = 1.0;
        return;
    }
    c[XK_NC_RPREV] = rel;
    c[XK_NC_RSTEPS] += 1.0;
}

__global__ void sv_refine_end_kernel(const double* __restrict__ c, const double* __restrict__ red,
                                     double* __restrict__ res_rel) {
    if (threadIdx.x != 0 || c[XK_NC_STATE] != XK_NC_RUNNING) return;
    if (res_rel) *res_rel = *red;
}

static unsigned sv_grid(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

#define SV_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            set_error(std::string("linear solve: ") + #expr + ": " + hipGetErrorString(_e)); \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

#define SV_RB(expr)                                                                        \
    do {                                                                                   \
        rocblas_status _s = (expr);                                                        \
        if (_s != rocblas_status_success) {                                                \
            set_error(std::string("linear solve: ") + #expr + ": " + rocblas_status_to_string(_s)); \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

template <class T>
static hipError_t sv_upload(DBuf& b, const std::vector<T>& v) {
    hipError_t e = b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

int Ctx::solve_buffer(int which, void* host, int64_t nbytes, int64_t* size) {
    if (which < 0 || which >= SV_COUNT || (int)sv.size() != SV_COUNT) {
        set_error("xk_solve_buffer: no such buffer (run a solve first)");
        return -1;
    }
    if (size) *size = (int64_t)sv[which].bytes;
    if (host && nbytes > 0) {
        SV_TRY(hipDeviceSynchronize());
        SV_TRY(hipMemcpy(host, sv[which].p, std::min<size_t>((size_t)nbytes, sv[which].bytes), hipMemcpyDeviceToHost));
    }
    return 0;
}

void Ctx::release_solver() {
    for (auto& b : sv) b.release();
    if (sv_handle) (void)rocblas_destroy_handle((rocblas_handle)sv_handle);
    sv_handle = nullptr;
    solve_ready = false;
}

// Per pattern: slots, contact nodes and their per-chain numbering, the
// components of the contact graph and the buffer sizes.
int Ctx::solve_setup() {
    sv.resize(SV_COUNT);
    const int ns = n_sinc;
    SV_TRY(sv[SV_SLIST].ensure(sizeof(int) * (size_t)std::max(ns, 1)));
    SV_TRY(sv[SV_SN].ensure(sizeof(int4) * (size_t)std::max(ns, 1)));
    if (ns > 0) {
        hipLaunchKernelGGL(sv_slist_kernel, dim3(sv_grid(n_pairs)), dim3(256), 0, 0, nullptr, n_pairs,
                           (const int*)P.sinc,
                           (int*)sv[SV_SLIST].p);
        SV_TRY(hipGetLastError());
        hipLaunchKernelGGL(sv_slot_nodes_kernel, dim3(sv_grid(ns)), dim3(256), 0, 0, nullptr, ns,
                           (const int*)sv[SV_SLIST].p,
                           P, (int4*)sv[SV_SN].p);
        SV_TRY(hipGetLastError());
    }
    std::vector<int4> sn(ns);
    if (ns) SV_TRY(hipMemcpy(sn.data(), sv[SV_SN].p, sizeof(int4) * ns, hipMemcpyDeviceToHost));
    // contact nodes, numbered in node order (so contiguous per chain)
    std::vector<int> cn(n_nodes, -1), cnode;
    for (const int4& s : sn)
        for (int v : {s.x, s.y, s.z, s.w}) cn[v] = 0;
    for (int n = 0; n < n_nodes; ++n)
        if (cn[n] == 0) {
            cn[n] = (int)cnode.size();
            cnode.push_back(n);
        }
    std::vector<int> ch_cn0(n_chains), ch_m(n_chains);
    std::vector<int64_t> ch_yoff(n_chains), ch_coff(n_chains);
    int64_t ytot = 0, ctot = 0;
    {
        int k = 0;
        for (int c = 0; c < n_chains; ++c) {
            const int n0 = cstart_h[c], n1 = cstart_h[c + 1];
            ch_cn0[c] = k;
            int cnt = 0;
            while (k < (int)cnode.size() && cnode[k] < n1) {
                ++k;
                ++cnt;
            }
            const int m = 3 * cnt;
            ch_m[c] = m;
            ch_yoff[c] = ytot;
            ch_coff[c] = ctot;
            ytot += (int64_t)(n1 - n0) * 6 * m;
            ctot += (int64_t)m * m;
        }
    }
    // components: chains joined by pairs (union-find), slots ordered by component
    std::vector<int> parent(n_chains);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (const int4& s : sn) {
        const int a = find(node_chain_h[s.x]), b = find(node_chain_h[s.z]);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
    std::vector<int> comp_of_root(n_chains, -1), slot_comp(ns);
    int ncomp = 0;
    for (int q = 0; q < ns; ++q) {
        const int r = find(node_chain_h[sn[q].x]);
        if (comp_of_root[r] < 0) comp_of_root[r] = ncomp++;
        slot_comp[q] = comp_of_root[r];
    }
    sv_comp_n.assign(ncomp, 0);
    for (int q = 0; q < ns; ++q) sv_comp_n[slot_comp[q]]++;
    sv_comp_pos.assign(ncomp + 1, 0);
    for (int c = 0; c < ncomp; ++c) sv_comp_pos[c + 1] = sv_comp_pos[c] + sv_comp_n[c];
    std::vector<int> pos(ns), cslot(ns);
    {
        std::vector<int64_t> fill(sv_comp_pos.begin(), sv_comp_pos.end() - 1);
        for (int q = 0; q < ns; ++q) {
            const int p = (int)fill[slot_comp[q]]++;
            pos[q] = p;
            cslot[p] = q;
        }
    }
    sv_comp_soff.assign(ncomp + 1, 0);
    for (int c = 0; c < ncomp; ++c) sv_comp_soff[c + 1] = sv_comp_soff[c] + 9 * (int64_t)sv_comp_n[c] * sv_comp_n[c];
    const int64_t stot = sv_comp_soff[ncomp];
    sv_comp_piv.assign(ncomp, 0);
    if ((double)stot * 8.0 > 64.0 * (1ull << 30)) {
        set_error("linear solve: the contact capacitance matrices exceed 64 GiB");
        return -1;
    }
    if (ns > 0 && *std::max_element(sv_comp_n.begin(), sv_comp_n.end()) > 65535) {
        set_error("linear solve: a contact component holds more than 65535 contributing pairs");
        return -1;
    }
    SV_TRY(sv_upload(sv[SV_CN], cn));
    SV_TRY(sv_upload(sv[SV_CNODE], cnode));
    SV_TRY(sv_upload(sv[SV_CHCN0], ch_cn0));
    SV_TRY(sv_upload(sv[SV_CHM], ch_m));
    SV_TRY(sv_upload(sv[SV_CHYOFF], ch_yoff));
    SV_TRY(sv_upload(sv[SV_CHCOFF], ch_coff));
    SV_TRY(sv_upload(sv[SV_POS], pos));
    SV_TRY(sv_upload(sv[SV_CSLOT], cslot));
    const size_t nd = 6 * (size_t)n_nodes;
    SV_TRY(sv[SV_KB].ensure((size_t)std::max(n_blocks, 1) * 9 * sizeof(double)));
    SV_TRY(sv[SV_AB].ensure(sizeof(double) * 108 * (size_t)n_nodes));
    SV_TRY(sv[SV_LF].ensure(sizeof(double) * 36 * (size_t)n_nodes));
    SV_TRY(sv[SV_SINV].ensure(sizeof(double) * 36 * (size_t)n_nodes));
    SV_TRY(sv[SV_SG].ensure(sizeof(double) * 9 * (size_t)std::max(ns, 1)));
    SV_TRY(sv[SV_SC].ensure(sizeof(double) * 4 * (size_t)std::max(ns, 1)));
    SV_TRY(sv[SV_Y].ensure(sizeof(double) * (size_t)std::max<int64_t>(ytot, 1)));
    SV_TRY(sv[SV_C].ensure(sizeof(double) * (size_t)std::max<int64_t>(ctot, 1)));
    SV_TRY(sv[SV_S].ensure(sizeof(double) * (size_t)std::max<int64_t>(stot, 1)));
    SV_TRY(sv[SV_IPIV].ensure(sizeof(int64_t) * 3 * (size_t)std::max(ns, 1)));
    SV_TRY(sv[SV_INFO].ensure(sizeof(int64_t) * (size_t)std::max(ncomp, 1)));
    SV_TRY(sv[SV_PIV32].ensure(sizeof(int) * (3 * (size_t)std::max(ns, 1) + (size_t)std::max(ncomp, 1))));
    SV_TRY(sv[SV_W].ensure(sizeof(double) * 3 * (size_t)std::max(ns, 1)));
    SV_TRY(sv[SV_W2].ensure(sizeof(double) * 3 * (size_t)std::max(ns, 1)));
    SV_TRY(sv[SV_PERM].ensure(sizeof(int) * 3 * (size_t)std::max(ns, 1)));
    for (int b : {SV_B, SV_X, SV_R, SV_T}) SV_TRY(sv[b].ensure(sizeof(double) * nd));
    SV_TRY(sv[SV_RED].ensure(64));
    if (!sv_handle) {
        rocblas_handle h = nullptr;
        SV_RB(rocblas_create_handle(&h));
        sv_handle = h;
    }
    solve_ready = true;
    sv_gepp = false;   // a new pattern starts with the LU without interchanges
    return 0;
}

// x <- A~^-1 r (Woodbury), r and x device vectors of 6 n_nodes; T scratch
int Ctx::solve_apply(const double* r, double* x, hipStream_t st, const double* stop) {
    const int ns = n_sinc;
    const int nd = 6 * n_nodes;
    double* T = (double*)sv[SV_T].p;
    double* W = (double*)sv[SV_W].p;
    const uint8_t* fixed = sv_fixed;
    // T = A~_b^-1 r (without contact the result itself)
    if (ns == 0) T = x;
    hipLaunchKernelGGL(sv_chain_solve_kernel, dim3(sv_grid(n_chains, 8)), dim3(64), 0, st, stop, n_chains,
                       (const int*)d_cstart.p, (const double*)sv[SV_AB].p, (const double*)sv[SV_LF].p,
                       (const double*)sv[SV_SINV].p, r, T);
    SV_TRY(hipGetLastError());
    if (ns == 0) return 0;
    hipLaunchKernelGGL(sv_project_kernel, dim3(sv_grid(ns)), dim3(256), 0, st, stop, ns, (const int*)sv[SV_POS].p,
                       (const int4*)sv[SV_SN].p, (const double*)sv[SV_SG].p, (const double*)sv[SV_SC].p, fixed,
                       (const double*)T, W);
    SV_TRY(hipGetLastError());
    double* W2 = (double*)sv[SV_W2].p;
    for (size_t c = 0; c < sv_comp_n.size(); ++c) {
        const int n = 3 * sv_comp_n[c];
        const int nb = (n + TB - 1) / TB;
        const double* LU = (const double*)sv[SV_S].p + sv_comp_soff[c];
        // permuted right-hand side in W2, the lower solve into W (its
        // projection is consumed), the upper solve back into W2
        double* y = W2 + 3 * sv_comp_pos[c];
        double* z = W + 3 * sv_comp_pos[c];
        hipLaunchKernelGGL(sv_permute_kernel, dim3(sv_grid(n)), dim3(256), 0, st, stop, n,
                           (const int*)sv[SV_PERM].p + 3 * sv_comp_pos[c], (const double*)z, y);
        SV_TRY(hipGetLastError());
        const int ts = sv_tri_steps(nb);
        for (int s = 0; s < nb; s += ts) {
            const int ks = std::min(ts, nb - s);
            hipLaunchKernelGGL(sv_lower_step_kernel, dim3(std::max(1, nb - s - ks)), dim3(TB), 0, st, stop, n, LU, y,
                               z, s, ks);
            SV_TRY(hipGetLastError());
        }
        for (int s = nb - 1; s >= 0; s -= ts) {
            const int ks = std::min(ts, s + 1);
            hipLaunchKernelGGL(sv_upper_step_kernel, dim3(std::max(1, s + 1 - ks)), dim3(TB), 0, st, stop, n, LU, z,
                               y, s, ks);
            SV_TRY(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(sv_lift_kernel, dim3(sv_grid(nd)), dim3(256), 0, st, stop, n_nodes, (const int*)d_node_pptr.p,
                       (const int*)d_node_psrc.p, (const int*)sv[SV_POS].p, (const double*)sv[SV_SC].p, fixed, r,
                       (const double*)W2, x);
    SV_TRY(hipGetLastError());
    hipLaunchKernelGGL(sv_chain_solve_kernel, dim3(sv_grid(n_chains, 8)), dim3(64), 0, st, stop, n_chains,
                       (const int*)d_cstart.p, (const double*)sv[SV_AB].p, (const double*)sv[SV_LF].p,
                       (const double*)sv[SV_SINV].p, x, x);
    SV_TRY(hipGetLastError());
    return 0;
}

// Every component's capacitance block assembled from the compliance columns
// and LU-factored in place: without row interchanges (gepp false, blocks up to
// LU_MAX rows; info[c] = first bad pivot, no flag raised: the caller decides)
// or by rocSOLVER's partially pivoted getrf (a zero pivot raises
// XK_FLAG_SINGULAR); then the pivots as one permutation per component.
// Stream-ordered (capturable) except for the host permutation of blocks
// beyond PERM_MAX rows.
constexpr int64_t LU_MAX = 32768;   // rocBLAS's 32-bit offsets: n * n < 2^31

int Ctx::sv_factor_capacitance(bool gepp, hipStream_t st) {
    const int ns = n_sinc;
    const uint8_t* fixed = sv_fixed;
    rocblas_handle h = (rocblas_handle)sv_handle;
    for (size_t c = 0; c < sv_comp_n.size(); ++c) {
        const int Pc = sv_comp_n[c];
        double* S = (double*)sv[SV_S].p + sv_comp_soff[c];
        hipLaunchKernelGGL(sv_capacitance_kernel, dim3(sv_grid(Pc, 64), Pc), dim3(64), 0, st, nstop, Pc,
                           (const int*)sv[SV_CSLOT].p + sv_comp_pos[c], (const int4*)sv[SV_SN].p,
                           (const double*)sv[SV_SG].p, (const double*)sv[SV_SC].p, (const int*)d_node_chain.p,
                           (const int*)sv[SV_CN].p, (const int*)sv[SV_CHCN0].p, (const int*)sv[SV_CHM].p,
                           (const int64_t*)sv[SV_CHCOFF].p, fixed, (const double*)sv[SV_C].p, S);
        SV_TRY(hipGetLastError());
        const int64_t n = 3 * (int64_t)Pc;
        int64_t* ipiv = (int64_t*)sv[SV_IPIV].p + 3 * sv_comp_pos[c];
        int64_t* info = (int64_t*)sv[SV_INFO].p + c;
        if (!gepp && n <= LU_MAX) {
            // Outer panels of LU_NP panels of LU_NB columns: each panel first takes
            // the outer panel's earlier panels on its own columns and rows (thin
            // DGEMMs), then is factored; one rank-(LU_NP LU_NB) DGEMM then updates the
            // trailing matrix with the whole outer panel, so the trailing matrix
            // streams through HBM once per LU_NP LU_NB columns.
            const double one = 1.0, minus_one = -1.0;
            const size_t ld = (size_t)n;
            auto panel = [&](int64_t j0) -> int {
                const int w = (int)std::min<int64_t>(LU_NB, n - j0);
                const int64_t rest = n - j0 - w;
                const int nrw = (int)((rest + LU_NB - 1) / LU_NB);
                if (w == LU_NB)
                    hipLaunchKernelGGL(sv_lu_panel_kernel<true>, dim3(std::max(1, 2 * nrw)), dim3(256), 0, st, nstop,
                                       (int)n, (int)j0, nrw, S, info);
                else
                    hipLaunchKernelGGL(sv_lu_panel_kernel<false>, dim3(std::max(1, 2 * nrw)), dim3(256), 0, st, nstop,
                                       (int)n, (int)j0, nrw, S, info);
                SV_TRY(hipGetLastError());
                return 0;
            };
            // C (m x q) -= A (m x k) B (k x q), all in S
            auto update = [&](int64_t m, int64_t q, int64_t k, const double* A_, const double* B_, double* C_) -> int {
                if (m <= 0 || q <= 0) return 0;
                SV_RB(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m,
                                    (rocblas_int)q, (rocblas_int)k, &minus_one, A_, (rocblas_int)n, B_,
                                    (rocblas_int)n, &one, C_, (rocblas_int)n));
                return 0;
            };
            for (int64_t J = 0; J < n; J += (int64_t)LU_NP * LU_NB) {
                const int64_t jE = std::min<int64_t>(n, J + (int64_t)LU_NP * LU_NB);
                for (int64_t jp = J; jp < jE; jp += LU_NB) {
                    const int64_t jp1 = std::min<int64_t>(jp + LU_NB, n);
                    if (jp > J) {
                        // the panel's columns (rows from jp down) and rows (columns right of
                        // it) take the outer panel's earlier panels, rank jp - J
                        const int64_t kk = jp - J;
                        if (update(n - jp, jp1 - jp, kk, S + jp + J * ld, S + J + jp * ld, S + jp + jp * ld) != 0)
                            return -1;
                        if (update(jp1 - jp, n - jp1, kk, S + jp + J * ld, S + J + jp1 * ld, S + jp + jp1 * ld) != 0)
                            return -1;
                    }
                    if (panel(jp) != 0) return -1;
                }
                if (jE >= n) break;
                // the trailing matrix, once per outer panel
                if (update(n - jE, n - jE, jE - J, S + jE + J * ld, S + J + jE * ld, S + jE + jE * ld) != 0)
                    return -1;
            }
            hipLaunchKernelGGL(sv_ipiv_identity_kernel, dim3(sv_grid(n)), dim3(256), 0, st, nstop, n, ipiv);
            SV_TRY(hipGetLastError());
            sv_comp_piv[c] = 0;
        } else {
            if (n <= 32768) {
                // 32-bit interface (its element offsets are 32-bit: n * n must stay far below
                // 2^31, hence n <= 32768), pivots and info widened to the 64-bit buffers afterwards
                int* piv32 = (int*)sv[SV_PIV32].p + 3 * sv_comp_pos[c];
                int* info32 = (int*)sv[SV_PIV32].p + 3 * (size_t)std::max(ns, 1) + c;
                SV_RB(rocsolver_dgetrf(h, (rocblas_int)n, (rocblas_int)n, S, (rocblas_int)n, piv32, info32));
                hipLaunchKernelGGL(sv_widen_kernel, dim3(sv_grid(n)), dim3(256), 0, st, nstop, n, (const int*)piv32,
                                   ipiv,
                                   (const int*)info32, info);
                SV_TRY(hipGetLastError());
            } else {
                SV_RB(rocsolver_dgetrf_64(h, n, n, S, n, ipiv, info));
            }
            hipLaunchKernelGGL(sv_info_kernel, dim3(1), dim3(64), 0, st, nstop, 1, (const int64_t*)info,
                               (int*)d_flags.p);
            SV_TRY(hipGetLastError());
            sv_comp_piv[c] = 1;
        }
    }
    // the row interchanges of every component as one permutation: on the
    // device when every block fits the permutation kernel, else on the host
    bool on_device = true;
    for (size_t c = 0; c < sv_comp_n.size(); ++c) on_device &= !sv_comp_piv[c] || 3 * sv_comp_n[c] <= PERM_MAX;
    if (on_device) {
        for (size_t c = 0; c < sv_comp_n.size(); ++c) {
            const int n = 3 * sv_comp_n[c];
            int* perm = (int*)sv[SV_PERM].p + 3 * sv_comp_pos[c];
            if (sv_comp_piv[c])
                hipLaunchKernelGGL(sv_perm_kernel, dim3(1), dim3(256), 0, st, nstop, n,
                                   (const int64_t*)sv[SV_IPIV].p + 3 * sv_comp_pos[c], perm, (int*)d_flags.p);
            else
                hipLaunchKernelGGL(sv_perm_identity_kernel, dim3(sv_grid(n)), dim3(256), 0, st, nstop, n, perm);
            SV_TRY(hipGetLastError());
        }
        return 0;
    }
    const size_t np3 = 3 * (size_t)ns;
    piv_h.resize(np3);
    perm_h.resize(np3);
    SV_TRY(hipMemcpyAsync(piv_h.data(), sv[SV_IPIV].p, sizeof(int64_t) * np3, hipMemcpyDeviceToHost, st));
    double state = 0.0;
    if (nstop) SV_TRY(hipMemcpyAsync(&state, nstop, sizeof(double), hipMemcpyDeviceToHost, st));
    SV_TRY(hipStreamSynchronize(st));
    // a solve enqueued past the end of a device-controlled call factored stale
    // scratch: its pivots mean nothing and nothing reads the permutation
    if (state != XK_NC_RUNNING) return 0;
    for (size_t c = 0; c < sv_comp_n.size(); ++c) {
        const int64_t o = 3 * sv_comp_pos[c], n = 3 * (int64_t)sv_comp_n[c];
        int* p = perm_h.data() + o;
        for (int64_t i = 0; i < n; ++i) p[i] = (int)i;
        for (int64_t k = 0; k < n; ++k) {
            const int64_t q = piv_h[o + k] - 1;
            if (q < 0 || q >= n) {
                set_error("linear solve: pivot index out of range");
                return -1;
            }
            std::swap(p[k], p[q]);
        }
    }
    SV_TRY(hipMemcpyAsync(sv[SV_PERM].p, perm_h.data(), sizeof(int) * np3, hipMemcpyHostToDevice, st));
    return 0;
}

int Ctx::linear_solve(const double* u, const uint8_t* fixed, double mass_coef, const double* M, const double* R,
                      double* du, int refine, double tol, double* res_rel, int* n_refine, hipStream_t st) {
    if (!direct) {
        set_error("linear solve: the device solve needs element chains (the direct element path)");
        return -1;
    }
    if (!pattern_valid || !tangent_staged) {
        set_error("linear solve: assemble the iteration with K on the current pattern first");
        return -1;
    }
    if (!solve_ready && solve_setup() != 0) return -1;
    sv_fixed = fixed;
    // with a device Newton control record attached and tol > 0 the refinement
    // stops on the device (no host read): each step is gated by a word of the
    // record that the step's norm test sets, so the steps after the stop are
    // no-ops; a solve left above tol is then the control's to hand to the host
    const bool dev_refine = nstop != nullptr && tol > 0.0;
    SV_RB(rocblas_set_stream((rocblas_handle)sv_handle, st));
    const int ns = n_sinc;
    const int nd = 6 * n_nodes;
    // beam-only tangent with the mass term, then the chain factors.  In both
    // precision forms the solve factors the FP64 element tangent and the FP64
    // pair records: the Newton step of a frictional strand is sensitive to
    // the FP32 rounding of the tangent (the 91-wire squeeze takes another
    // trajectory with it), so the mixed form's FP32 K is an output for
    // consumers of the assembled matrix, not the operator this solve inverts.
    SV_TRY(launch_elements(u, false, sv[SV_KB].p, mass_coef, M, st, 0));
    hipLaunchKernelGGL(sv_extract_kernel<double>, dim3(sv_grid(nd)), dim3(256), 0, st, nstop, n_nodes,
                       (const int8_t*)d_nflag.p, (const int*)d_gptr.p, (const int*)d_gcol.p,
                       (const double*)sv[SV_KB].p, fixed, (double*)sv[SV_AB].p);
    SV_TRY(hipGetLastError());
    hipLaunchKernelGGL(sv_factor_kernel, dim3(n_chains), dim3(64), 0, st, nstop, (const int*)d_cstart.p,
                       (const double*)sv[SV_AB].p, (double*)sv[SV_LF].p, (double*)sv[SV_SINV].p, (int*)d_flags.p);
    SV_TRY(hipGetLastError());
    // capacitance LU: without row interchanges unless partial pivoting was
    // asked for or this pattern has already needed it
    bool gepp = sv_gepp || sv_lu_kind == 1;
    if (ns > 0) {
        hipLaunchKernelGGL(sv_slot_pack_kernel<double>, dim3(sv_grid(ns)), dim3(256), 0, st, nstop, ns,
                           (const double*)d_sg.p, (double*)sv[SV_SG].p, (double*)sv[SV_SC].p);
        SV_TRY(hipGetLastError());
        hipLaunchKernelGGL(sv_compliance_kernel, dim3(n_chains), dim3(256), 0, st, nstop, (const int*)d_cstart.p,
                           (const int*)sv[SV_CN].p, (const int*)sv[SV_CNODE].p, (const int*)sv[SV_CHCN0].p,
                           (const int*)sv[SV_CHM].p, (const int64_t*)sv[SV_CHYOFF].p,
                           (const int64_t*)sv[SV_CHCOFF].p, fixed, (const double*)sv[SV_AB].p,
                           (const double*)sv[SV_LF].p, (const double*)sv[SV_SINV].p, (double*)sv[SV_Y].p,
                           (double*)sv[SV_C].p);
        SV_TRY(hipGetLastError());
        if (sv_factor_capacitance(gepp, st) != 0) return -1;
        if (!gepp && tol > 0.0 && !dev_refine) {
            // eager solve: a zero or non-finite pivot of the LU without
            // interchanges is read back at once and the blocks factored again
            // with partial pivoting
            const size_t nc = sv_comp_n.size();
            info_h.resize(nc);
            SV_TRY(hipMemcpyAsync(info_h.data(), sv[SV_INFO].p, sizeof(int64_t) * nc, hipMemcpyDeviceToHost, st));
            SV_TRY(hipStreamSynchronize(st));
            bool bad = false;
            for (int64_t v : info_h) bad |= v != 0;
            if (bad) {
                gepp = sv_gepp = true;
                ++sv_fallbacks;
                if (sv_factor_capacitance(true, st) != 0) return -1;
            }
        }
    }
    double* b = (double*)sv[SV_B].p;
    double* r = (double*)sv[SV_R].p;
    double* x = (double*)sv[SV_X].p;
    hipLaunchKernelGGL(sv_rhs_kernel, dim3(sv_grid(nd)), dim3(256), 0, st, nstop, nd, R, fixed, b);
    SV_TRY(hipGetLastError());
    if (solve_apply(b, du, st, nstop) != 0) return -1;
    // iterative refinement against the assembled K: a fixed number of steps
    // (tol <= 0, stream-ordered) or until ||b - A du|| <= tol ||b|| (one
    // host read of the norm per step)
    // residual of the operator the solve inverts, A~_b + U~ G U~^T (in the
    // FP64 form equal to the assembled K up to rounding; in the mixed form K
    // accumulates its contact blocks in FP32 and differs by that rounding)
    const double* gate = dev_refine ? nstop + XK_NC_RGATE : nstop;
    auto residual = [&]() -> int {
        // b - A~_b du into x (with contact terms to subtract) or straight into r
        hipLaunchKernelGGL(sv_chain_residual_kernel, dim3(sv_grid(nd)), dim3(256), 0, st, gate, n_nodes,
                           (const double*)sv[SV_AB].p, (const double*)b, (const double*)du, ns > 0 ? x : r);
        SV_TRY(hipGetLastError());
        if (ns > 0) {
            hipLaunchKernelGGL(sv_project_kernel, dim3(sv_grid(ns)), dim3(256), 0, st, gate, ns,
                               (const int*)sv[SV_POS].p,
                               (const int4*)sv[SV_SN].p, (const double*)sv[SV_SG].p, (const double*)sv[SV_SC].p,
                               fixed, (const double*)du, (double*)sv[SV_W].p);
            SV_TRY(hipGetLastError());
            hipLaunchKernelGGL(sv_lift_kernel, dim3(sv_grid(nd)), dim3(256), 0, st, gate, n_nodes,
                               (const int*)d_node_pptr.p,
                               (const int*)d_node_psrc.p, (const int*)sv[SV_POS].p, (const double*)sv[SV_SC].p,
                               fixed, (const double*)x, (const double*)sv[SV_W].p, r);
        }
        SV_TRY(hipGetLastError());
        return 0;
    };
    double* red = (double*)sv[SV_RED].p;
    int steps = 0;
    bool have_r = false;
    double rel_last = HUGE_VAL;   // last relative residual read on the host (tol > 0)
    auto refine_loop = [&]() -> int {
        have_r = false;
        double prev = HUGE_VAL;
        if (dev_refine) {
            hipLaunchKernelGGL(sv_refine_begin_kernel, dim3(1), dim3(64), 0, st, nstop);
            SV_TRY(hipGetLastError());
        }
        for (int k = 0; k < refine; ++k) {
            if (residual() != 0) return -1;
            if (dev_refine) {
                hipLaunchKernelGGL(sv_relnorm_kernel, dim3(1), dim3(1024), 0, st, gate, nd, (const double*)r,
                                   (const double*)b, red);
                hipLaunchKernelGGL(sv_refine_test_kernel, dim3(1), dim3(64), 0, st, nstop, (const double*)red, tol);
                SV_TRY(hipGetLastError());
            } else if (tol > 0.0) {
                double rel = 0.0;
                hipLaunchKernelGGL(sv_relnorm_kernel, dim3(1), dim3(1024), 0, st, nstop, nd, (const double*)r,
                                   (const double*)b, red);
                SV_TRY(hipGetLastError());
                SV_TRY(hipMemcpyAsync(&rel, red, sizeof(double), hipMemcpyDeviceToHost, st));
                SV_TRY(hipStreamSynchronize(st));
                rel_last = rel;
                // done, at the rounding floor (a step no longer halves the
                // residual), or not finite
                if (rel <= tol || rel > 0.5 * prev || !(rel <= DBL_MAX)) {
                    have_r = true;
                    return 0;
                }
                prev = rel;
            }
            if (solve_apply(r, x, st, gate) != 0) return -1;
            hipLaunchKernelGGL(sv_axpy_kernel, dim3(sv_grid(nd)), dim3(256), 0, st, gate, nd, (const double*)x, du);
            SV_TRY(hipGetLastError());
            ++steps;
        }
        if (dev_refine) {
            // all steps taken without a stop: the residual of the last one
            if (residual() != 0) return -1;
            hipLaunchKernelGGL(sv_relnorm_kernel, dim3(1), dim3(1024), 0, st, gate, nd, (const double*)r,
                               (const double*)b, red);
            SV_TRY(hipGetLastError());
            have_r = true;
        } else if (tol > 0.0) {
            if (residual() != 0) return -1;
            hipLaunchKernelGGL(sv_relnorm_kernel, dim3(1), dim3(1024), 0, st, nstop, nd, (const double*)r,
                               (const double*)b, red);
            SV_TRY(hipGetLastError());
            SV_TRY(hipMemcpyAsync(&rel_last, red, sizeof(double), hipMemcpyDeviceToHost, st));
            SV_TRY(hipStreamSynchronize(st));
            have_r = true;
        }
        return 0;
    };
    if (refine_loop() != 0) return -1;
    if (ns > 0 && !gepp && tol > 0.0 && !dev_refine && !(rel_last <= tol)) {
        // the LU without interchanges left the refinement above tol: factor
        // with partial pivoting (kept for the rest of this pattern) and solve again
        gepp = sv_gepp = true;
        ++sv_fallbacks;
        if (sv_factor_capacitance(true, st) != 0) return -1;
        if (solve_apply(b, du, st, nstop) != 0) return -1;
        if (refine_loop() != 0) return -1;
    }
    if (dev_refine) {
        // the relative residual where the refinement stopped (red) -> res_rel
        hipLaunchKernelGGL(sv_refine_end_kernel, dim3(1), dim3(64), 0, st, nstop, (const double*)red, res_rel);
        SV_TRY(hipGetLastError());
        if (n_refine) *n_refine = -1;   // decided on the device: XK_NC_RSTEPS
        return 0;
    }
    if (n_refine) *n_refine = steps;
    if (res_rel) {
        if (!have_r && residual() != 0) return -1;
        hipLaunchKernelGGL(sv_relnorm_kernel, dim3(1), dim3(1024), 0, st, nstop, nd, (const double*)r, (const double*)b,
                           res_rel);
        SV_TRY(hipGetLastError());
    }
    return 0;
}

}  // namespace xk
