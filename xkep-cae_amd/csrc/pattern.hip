// Tangent pattern of an increment, built on the GPU (direct element path).
//
// The pattern is the union of the beam blocks and the blocks of the pairs that
// can contribute within the increment (status != INACTIVE or p_n > 0, with the
// coating on also a compression > 0), as a
// 3x3-block CSR over DOF groups (2n translations of node n, 2n+1 rotations).
// On element chains with consecutive node numbers the beam blocks of the row
// of node n are implied: column groups of n-1 (if an element joins n-1 and n),
// of n (if n has an element) and of n+1 (if an element joins n and n+1).  Only
// the contact blocks -- translational blocks between nodes of the two segments
// of a contributing pair that no element couples -- are sorted (radix sort of
// (row, column) keys) and merged row by row into the implied beam columns.
// Element and pair contributions are then located by a short search within
// their row, and the per-block and per-node contribution lists are ordered by
// a stable key sort, (block, code) and (node, code), so every list holds its
// contributions in element / pair order as the host builder does: results are
// bitwise those of the host-built pattern.

#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

namespace xk {

__device__ __forceinline__ bool node_has(const int8_t* nflag, int n) { return nflag[n] & 1; }
__device__ __forceinline__ bool node_link(const int8_t* nflag, int n, int n_nodes) {
    return n >= 0 && n < n_nodes && (nflag[n] & 2);   // an element joins n and n + 1
}
// translational blocks (nk, nl) that the beam part already holds
__device__ __forceinline__ bool beam_coupled(const int8_t* nflag, int n_nodes, int a, int b) {
    if (a == b) return node_has(nflag, a);
    if (b == a + 1) return node_link(nflag, a, n_nodes);
    if (a == b + 1) return node_link(nflag, b, n_nodes);
    return false;
}
__device__ __forceinline__ int pair_node(const PairSoA& P, int64_t i, int k) {
    return k == 0 ? P.na0[i] : (k == 1 ? P.na1[i] : (k == 2 ? P.nb0[i] : P.nb1[i]));
}

// contributing: status != INACTIVE or p_n > 0; with the coating on (cc != nullptr)
// also a compression (current, previous step or probed) > 0
__global__ void pat_contrib_kernel(int64_t n, const int8_t* __restrict__ status, const double* __restrict__ p_n,
                                   const double* __restrict__ cc, const double* __restrict__ ccp,
                                   const int8_t* __restrict__ cneed, int64_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool k = status[i] != ST_INACTIVE || p_n[i] > 0.0;
    if (cc) k = k || cc[i] > 0.0 || ccp[i] > 0.0 || cneed[i] != 0;
    keep[i] = k ? 1 : 0;
}

__global__ void pat_sinc_kernel(int64_t n, const int64_t* __restrict__ keep, const int64_t* __restrict__ pos,
                                int* __restrict__ sinc, int* __restrict__ slist) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keep[i]) {
        sinc[i] = (int)pos[i];
        slist[pos[i]] = (int)i;
    } else {
        sinc[i] = -1;
    }
}

__global__ void pat_contact_keys_kernel(int64_t ns, const int* __restrict__ slist, PairSoA P,
                                        const int8_t* __restrict__ nflag, int n_nodes, uint64_t* __restrict__ keys) {
    // blocks the beam part holds get the key of row 2 n_nodes, past every real row
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 16 * ns) return;
    const int64_t i = slist[t >> 4];
    const int k = (int)((t >> 2) & 3), l = (int)(t & 3);
    const int a = pair_node(P, i, k), b = pair_node(P, i, l);
    keys[t] = beam_coupled(nflag, n_nodes, a, b) ? ((uint64_t)(2 * n_nodes) << 32)
                                                 : (((uint64_t)(2 * a) << 32) | (uint64_t)(2 * b));
}

__global__ void pat_unique_flag_kernel(int64_t m, const uint64_t* __restrict__ s, uint64_t beam_key,
                                       int64_t* __restrict__ flag) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) flag[j] = (s[j] != beam_key && (j == 0 || s[j] != s[j - 1])) ? 1 : 0;
}

__global__ void pat_unique_write_kernel(int64_t m, const uint64_t* __restrict__ s, const int64_t* __restrict__ flag,
                                        const int64_t* __restrict__ upos, uint64_t* __restrict__ ukeys,
                                        int* __restrict__ ccount) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m || !flag[j]) return;
    ukeys[upos[j]] = s[j];
    atomicAdd(&ccount[(int)(s[j] >> 32)], 1);
}

// number of beam column groups of the row of node n (each neighbour node: 2 groups)
__device__ __forceinline__ int beam_cols(const int8_t* nflag, int n_nodes, int n, int* cols) {
    int m = 0;
    if (node_link(nflag, n - 1, n_nodes)) { cols[m++] = 2 * (n - 1); cols[m++] = 2 * (n - 1) + 1; }
    if (node_has(nflag, n)) { cols[m++] = 2 * n; cols[m++] = 2 * n + 1; }
    if (node_link(nflag, n, n_nodes)) { cols[m++] = 2 * (n + 1); cols[m++] = 2 * (n + 1) + 1; }
    return m;
}

__global__ void pat_row_count_kernel(int ng, const int8_t* __restrict__ nflag, int n_nodes,
                                     const int* __restrict__ ccount, int64_t* __restrict__ gcount,
                                     int64_t* __restrict__ ccount64) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int cols[6];
    gcount[g] = beam_cols(nflag, n_nodes, g >> 1, cols) + ccount[g];
    ccount64[g] = ccount[g];
}

__global__ void pat_row_fill_kernel(int ng, const int8_t* __restrict__ nflag, int n_nodes,
                                    const int64_t* __restrict__ gstart, const int64_t* __restrict__ cstart,
                                    const int* __restrict__ ccount, const uint64_t* __restrict__ ukeys,
                                    int* __restrict__ gptr, int* __restrict__ gcol, int* __restrict__ brow,
                                    int8_t* __restrict__ bbeam) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int cols[6];
    const int nb = beam_cols(nflag, n_nodes, g >> 1, cols);
    const int nc = ccount[g];
    const uint64_t* ck = ukeys + cstart[g];
    int64_t o = gstart[g];
    gptr[g] = (int)o;
    int ib = 0, ic = 0;
    while (ib < nb || ic < nc) {
        const int cb = ib < nb ? cols[ib] : 0x7fffffff;
        const int cc = ic < nc ? (int)(ck[ic] & 0xffffffffu) : 0x7fffffff;
        const bool take_b = cb < cc;
        gcol[o] = take_b ? cb : cc;
        brow[o] = g;
        bbeam[o] = take_b ? 1 : 0;
        if (take_b) ++ib;
        else ++ic;
        ++o;
    }
}

__global__ void pat_gptr_end_kernel(int ng, int64_t total, int* __restrict__ gptr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) gptr[ng] = (int)total;
}

__device__ __forceinline__ int find_block(const int* __restrict__ gptr, const int* __restrict__ gcol, int row,
                                          int col) {
    int lo = gptr[row], hi = gptr[row + 1];
    while (lo < hi) {   // rows hold a handful of blocks: binary search in few steps
        const int mid = (lo + hi) >> 1;
        if (gcol[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void pat_eslot_kernel(int ne, const int* __restrict__ conn, const int* __restrict__ gptr,
                                 const int* __restrict__ gcol, int* __restrict__ eslot) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int n1 = conn[2 * e], n2 = conn[2 * e + 1];
    const int grp[4] = {2 * n1, 2 * n1 + 1, 2 * n2, 2 * n2 + 1};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) eslot[16 * (size_t)e + 4 * a + b] = find_block(gptr, gcol, grp[a], grp[b]);
}

__global__ void pat_pair_keys_kernel(int64_t ns, const int* __restrict__ slist, PairSoA P,
                                     const int* __restrict__ gptr, const int* __restrict__ gcol,
                                     uint64_t* __restrict__ keys2, uint64_t* __restrict__ keys3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 16 * ns) return;
    const int64_t q = t >> 4;
    const int64_t i = slist[q];
    const int k = (int)((t >> 2) & 3), l = (int)(t & 3);
    const int a = pair_node(P, i, k), b = pair_node(P, i, l);
    const int blk = find_block(gptr, gcol, 2 * a, 2 * b);
    keys2[t] = ((uint64_t)blk << 32) | (uint64_t)(q * 16 + k * 4 + l);
    if (l == 0) keys3[q * 4 + k] = ((uint64_t)a << 32) | (uint64_t)(q * 4 + k);
}

// sorted (head << 32 | code) keys -> counts per head and the codes in order
__global__ void pat_list_kernel(int64_t m, const uint64_t* __restrict__ s, int64_t* __restrict__ cnt,
                                int* __restrict__ src) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    src[j] = (int)(s[j] & 0xffffffffu);
    atomicAdd((unsigned long long*)&cnt[(int64_t)(s[j] >> 32)], 1ull);
}

// qpos[src[q]] = q: where each (slot, k) contribution sits in the node lists
__global__ void pat_inverse_kernel(int64_t m, const int* __restrict__ src, int* __restrict__ qpos) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m) qpos[src[q]] = (int)q;
}

__global__ void pat_i64_to_i32_kernel(int64_t n, const int64_t* __restrict__ a, int64_t last, int* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int)a[i];
    if (i == n) b[n] = (int)last;
}

__global__ void pat_cflag_kernel(int64_t nb, const int64_t* __restrict__ bcnt, int64_t* __restrict__ flag) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) flag[b] = bcnt[b] > 0 ? 1 : 0;
}

__global__ void pat_cdesc_kernel(int64_t nb, const int64_t* __restrict__ flag, const int64_t* __restrict__ fpos,
                                 const int* __restrict__ gptr, const int* __restrict__ brow,
                                 const int8_t* __restrict__ bbeam, const int* __restrict__ pptr,
                                 int* __restrict__ desc) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !flag[b]) return;
    const int r = brow[b], r0 = gptr[r], deg = gptr[r + 1] - r0;
    int* d = desc + 4 * fpos[b];
    d[0] = 9 * r0 + 3 * ((int)b - r0);
    d[1] = 3 * deg;
    d[2] = pptr[b];
    d[3] = pptr[b + 1] | (bbeam[b] ? (int)0x80000000u : 0);
}

static hipError_t sort_u64(Ctx* c, const uint64_t* in, uint64_t* out, int64_t n, int end_bit, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    size_t tmp = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, tmp, in, out, (size_t)n, 0, end_bit, st);
    if (e != hipSuccess) return e;
    if ((e = c->pat_tmp.ensure(std::max<size_t>(tmp, 1))) != hipSuccess) return e;
    return rocprim::radix_sort_keys(c->pat_tmp.p, tmp, in, out, (size_t)n, 0, end_bit, st);
}

static unsigned grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

static int bits_of(uint64_t v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return std::max(b, 1);
}

#define PT_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            set_error(std::string("pattern: ") + #expr + ": " + hipGetErrorString(_e));  \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

int Ctx::build_pattern_gpu(hipStream_t st) {
    const int64_t np = n_pairs;
    const int ng = 2 * n_nodes;
    std::vector<DBuf>& S = pat;
    S.resize(12);
    // 1. contributing pairs -> sinc, slist
    int64_t ns = 0;
    PT_TRY(S[0].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(np, 1)));   // keep
    PT_TRY(S[1].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(np, 1)));   // pos
    PT_TRY(S[2].ensure(sizeof(int) * (size_t)std::max<int64_t>(np, 1)));       // slist
    int64_t* keep = (int64_t*)S[0].p;
    int64_t* pos = (int64_t*)S[1].p;
    int* slist = (int*)S[2].p;
    if (np > 0) {
        hipLaunchKernelGGL(pat_contrib_kernel, dim3(grid(np)), dim3(256), 0, st, np, P.status, P.p_n,
                           coat_on ? (const double*)P.cc : nullptr, (const double*)P.ccp, (const int8_t*)P.cneed,
                           keep);
        PT_TRY(hipGetLastError());
        if (scan_exclusive_i64(this, 0, keep, pos, np, &ns, st) != 0) return -1;
        hipLaunchKernelGGL(pat_sinc_kernel, dim3(grid(np)), dim3(256), 0, st, np, keep, pos, P.sinc, slist);
        PT_TRY(hipGetLastError());
    }
    // 2. contact-only block keys, sorted, unique, counted per row
    const int64_t m = 16 * ns;
    if (m >= INT32_MAX) {
        set_error("pattern: too many contributing pairs");
        return -1;
    }
    PT_TRY(S[3].ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(m, 1)));   // keys
    PT_TRY(S[4].ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(m, 1)));   // sorted / ukeys
    PT_TRY(S[5].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(m, 1)));    // flag
    PT_TRY(S[6].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(m, 1)));    // upos
    PT_TRY(S[7].ensure(sizeof(int) * (size_t)ng + sizeof(int64_t) * 3 * ((size_t)ng + 1)));
    uint64_t* keys = (uint64_t*)S[3].p;
    uint64_t* sorted = (uint64_t*)S[4].p;
    int64_t* flag = (int64_t*)S[5].p;
    int64_t* upos = (int64_t*)S[6].p;
    int* ccount = (int*)S[7].p;
    int64_t* gcount = (int64_t*)((char*)S[7].p + sizeof(int) * (size_t)ng);
    int64_t* ccount64 = gcount + (ng + 1);
    int64_t* cstart = ccount64 + (ng + 1);
    PT_TRY(launch_fill_i32((size_t)ng, ccount, 0, st));
    int64_t n_uc = 0;
    if (m > 0) {
        hipLaunchKernelGGL(pat_contact_keys_kernel, dim3(grid(m)), dim3(256), 0, st, ns, slist, P,
                           (const int8_t*)d_nflag.p, n_nodes, keys);
        PT_TRY(hipGetLastError());
        PT_TRY(sort_u64(this, keys, sorted, m, 32 + bits_of((uint64_t)ng), st));
        hipLaunchKernelGGL(pat_unique_flag_kernel, dim3(grid(m)), dim3(256), 0, st, m, sorted,
                           (uint64_t)ng << 32, flag);
        PT_TRY(hipGetLastError());
        if (scan_exclusive_i64(this, 0, flag, upos, m, &n_uc, st) != 0) return -1;
        hipLaunchKernelGGL(pat_unique_write_kernel, dim3(grid(m)), dim3(256), 0, st, m, sorted, flag, upos, keys,
                           ccount);
        PT_TRY(hipGetLastError());
    }
    const uint64_t* ukeys = keys;   // unique contact keys, sorted by (row, column)
    // 3. rows: beam columns + contact columns
    hipLaunchKernelGGL(pat_row_count_kernel, dim3(grid(ng)), dim3(256), 0, st, ng, (const int8_t*)d_nflag.p, n_nodes,
                       ccount, gcount, ccount64);
    PT_TRY(hipGetLastError());
    int64_t nblk = 0, ncs = 0;
    PT_TRY(S[8].ensure(sizeof(int64_t) * ((size_t)ng + 1)));
    int64_t* gstart = (int64_t*)S[8].p;
    if (scan_exclusive_i64(this, 0, gcount, gstart, ng, &nblk, st) != 0) return -1;
    if (scan_exclusive_i64(this, 0, ccount64, cstart, ng, &ncs, st) != 0) return -1;
    if (9 * nblk + 64 > INT32_MAX) {
        set_error("tangent pattern exceeds 2^31 entries");
        return -1;
    }
    PT_TRY(d_gptr.ensure(sizeof(int) * ((size_t)ng + 1)));
    PT_TRY(d_gcol.ensure(sizeof(int) * (size_t)std::max<int64_t>(nblk, 1)));
    PT_TRY(d_brow.ensure(sizeof(int) * (size_t)std::max<int64_t>(nblk, 1)));
    PT_TRY(d_bbeam.ensure((size_t)std::max<int64_t>(nblk, 1)));
    hipLaunchKernelGGL(pat_row_fill_kernel, dim3(grid(ng)), dim3(256), 0, st, ng, (const int8_t*)d_nflag.p,
                       n_nodes, gstart, cstart, ccount, ukeys, (int*)d_gptr.p, (int*)d_gcol.p, (int*)d_brow.p,
                       (int8_t*)d_bbeam.p);
    PT_TRY(hipGetLastError());
    hipLaunchKernelGGL(pat_gptr_end_kernel, dim3(1), dim3(1), 0, st, ng, nblk, (int*)d_gptr.p);
    PT_TRY(hipGetLastError());
    // 4. element slots
    PT_TRY(d_eslot.ensure(sizeof(int) * 16 * (size_t)std::max(n_elems, 1)));
    if (n_elems > 0) {
        hipLaunchKernelGGL(pat_eslot_kernel, dim3(grid(n_elems)), dim3(256), 0, st, n_elems, (const int*)d_conn.p,
                           (const int*)d_gptr.p, (const int*)d_gcol.p, (int*)d_eslot.p);
        PT_TRY(hipGetLastError());
    }
    // 5. pair contributions per block (stable in pair order) and per node
    PT_TRY(S[9].ensure(sizeof(uint64_t) * (size_t)std::max<int64_t>(4 * ns, 1)));   // node keys
    PT_TRY(S[10].ensure(sizeof(int64_t) * ((size_t)std::max<int64_t>(nblk, (int64_t)n_nodes) + 1)));
    PT_TRY(S[11].ensure(sizeof(int64_t) * ((size_t)std::max<int64_t>(nblk, (int64_t)n_nodes) + 1)));
    uint64_t* keys3 = (uint64_t*)S[9].p;
    int64_t* cnt = (int64_t*)S[10].p;
    int64_t* cpos = (int64_t*)S[11].p;
    PT_TRY(d_blk_pptr.ensure(sizeof(int) * ((size_t)nblk + 1)));
    PT_TRY(d_blk_psrc.ensure(sizeof(int) * (size_t)std::max<int64_t>(m, 1)));
    PT_TRY(d_node_pptr.ensure(sizeof(int) * ((size_t)n_nodes + 1)));
    PT_TRY(d_node_psrc.ensure(sizeof(int) * (size_t)std::max<int64_t>(4 * ns, 1)));
    if (m > 0) {
        hipLaunchKernelGGL(pat_pair_keys_kernel, dim3(grid(m)), dim3(256), 0, st, ns, slist, P, (const int*)d_gptr.p,
                           (const int*)d_gcol.p, keys, keys3);
        PT_TRY(hipGetLastError());
        PT_TRY(sort_u64(this, keys, sorted, m, 32 + bits_of((uint64_t)nblk), st));
    }
    PT_TRY(hipMemsetAsync(cnt, 0, sizeof(int64_t) * ((size_t)nblk + 1), st));
    if (m > 0) {
        hipLaunchKernelGGL(pat_list_kernel, dim3(grid(m)), dim3(256), 0, st, m, sorted, cnt, (int*)d_blk_psrc.p);
        PT_TRY(hipGetLastError());
    }
    int64_t tot = 0;
    if (scan_exclusive_i64(this, 0, cnt, cpos, nblk, &tot, st) != 0) return -1;
    hipLaunchKernelGGL(pat_i64_to_i32_kernel, dim3(grid(nblk + 1)), dim3(256), 0, st, nblk, cpos, tot,
                       (int*)d_blk_pptr.p);
    PT_TRY(hipGetLastError());
    // contact descriptors: blocks with pair contributions, in block order
    int64_t nd = 0;
    PT_TRY(S[5].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(nblk, 1)));
    PT_TRY(S[6].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(nblk, 1)));
    flag = (int64_t*)S[5].p;
    upos = (int64_t*)S[6].p;
    hipLaunchKernelGGL(pat_cflag_kernel, dim3(grid(nblk)), dim3(256), 0, st, nblk, cnt, flag);
    PT_TRY(hipGetLastError());
    if (scan_exclusive_i64(this, 0, flag, upos, nblk, &nd, st) != 0) return -1;
    PT_TRY(d_cdesc.ensure(sizeof(int) * 4 * (size_t)std::max<int64_t>(nd, 1)));
    hipLaunchKernelGGL(pat_cdesc_kernel, dim3(grid(nblk)), dim3(256), 0, st, nblk, flag, upos, (const int*)d_gptr.p,
                       (const int*)d_brow.p, (const int8_t*)d_bbeam.p, (const int*)d_blk_pptr.p, (int*)d_cdesc.p);
    PT_TRY(hipGetLastError());
    // node lists
    if (ns > 0) PT_TRY(sort_u64(this, keys3, sorted, 4 * ns, 32 + bits_of((uint64_t)n_nodes), st));
    PT_TRY(hipMemsetAsync(cnt, 0, sizeof(int64_t) * ((size_t)n_nodes + 1), st));
    PT_TRY(d_node_qpos.ensure(sizeof(int) * (size_t)std::max<int64_t>(4 * ns, 1)));
    if (ns > 0) {
        hipLaunchKernelGGL(pat_list_kernel, dim3(grid(4 * ns)), dim3(256), 0, st, 4 * ns, sorted, cnt,
                           (int*)d_node_psrc.p);
        PT_TRY(hipGetLastError());
        hipLaunchKernelGGL(pat_inverse_kernel, dim3(grid(4 * ns)), dim3(256), 0, st, 4 * ns,
                           (const int*)d_node_psrc.p, (int*)d_node_qpos.p);
        PT_TRY(hipGetLastError());
    }
    if (scan_exclusive_i64(this, 0, cnt, cpos, n_nodes, &tot, st) != 0) return -1;
    hipLaunchKernelGGL(pat_i64_to_i32_kernel, dim3(grid(n_nodes + 1)), dim3(256), 0, st, (int64_t)n_nodes, cpos, tot,
                       (int*)d_node_pptr.p);
    PT_TRY(hipGetLastError());
    PT_TRY(hipStreamSynchronize(st));
    n_cdesc = (int)nd;
    n_sinc = (int)ns;
    n_blocks = (int)nblk;
    host_pattern_valid = false;
    return 0;
}

}  // namespace xk
