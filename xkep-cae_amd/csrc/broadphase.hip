// GPU broadphase and candidate management (DetectCandidatesProcess).
//
// Output contract: the exact set of element pairs (i < j) whose FP64
// axis-aligned boxes
//     lo = min(x0, x1) - (r + margin),  hi = max(x0, x1) + (r + margin)
// overlap inclusively on all three axes -- the set the reference's
// _broadphase_aabb returns (xkep_cae/contact/_broadphase.py:13-96), here in
// sorted (i, j) order.
//
// Method: a dense uniform grid.  One reduction pass gives the boxes' bounds,
// extents and a count of non-finite boxes (a diverged displacement is refused
// before any cell loop runs); the host sizes the grid from them.  Every box is
// binned into the cells its floor(lo * inv) .. floor(hi * inv) range covers
// (counting sort by cell).  One wavefront per cell then tests the pairs of its
// members and keeps a pair only in the cell holding floor(max(lo_i, lo_j) *
// inv), the one cell both boxes are guaranteed to share when they overlap --
// so every pair is found exactly once and the cell size only changes the
// work, never the set.  The pairs' keys are radix-sorted, which makes the list
// canonical.
//
// Filters of DetectCandidatesProcess (contact/_manager_process.py:200-217):
// shared node, optional same-layer exclusion, and the midpoint prescreen
// max(0, |m_a - m_b| - L_a/2 - L_b/2) <= r_a + r_b + extra (:259-286).
// Candidates then update the pair list: existing pairs outside the set become
// INACTIVE (state kept), new ones are appended with a fresh state (:219-250).
//
// Compiled with FP contraction off: box bounds, cell indices and the midpoint
// cutoff use the same roundings as the FP64 reference formulation.

#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"
#include "../../include/xkamd.h"

namespace xk {

// ------------------------------------------------------------- exclusive scan

constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_TILE = SCAN_ITEMS * SCAN_BLOCK;

__global__ void __launch_bounds__(SCAN_BLOCK)
scan_tile_kernel(int64_t n, const int64_t* __restrict__ in, int64_t* __restrict__ out, int64_t* __restrict__ sums) {
    __shared__ int64_t sh[SCAN_BLOCK];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    int64_t tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        tot += v[k];
    }
    sh[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        int64_t add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t run = sh[threadIdx.x] - tot;  // exclusive prefix of this lane
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BLOCK - 1 && sums) sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ void scan_add_kernel(int64_t n, int64_t* __restrict__ out, const int64_t* __restrict__ offs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / SCAN_TILE];
}


static int scan_rec(Ctx* c, int slot, const int64_t* in, int64_t* out, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (slot >= 6) {
        set_error("scan: input too large");
        return -1;
    }
    if (c->bp[slot].ensure(sizeof(int64_t) * (size_t)(2 * nb + 2)) != hipSuccess) {
        set_error("scan: out of memory");
        return -1;
    }
    int64_t* sums = (int64_t*)c->bp[slot].p;
    int64_t* soff = sums + nb + 1;
    hipLaunchKernelGGL(scan_tile_kernel, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, n, in, out, sums);
    if (nb > 1) {
        if (scan_rec(c, slot + 1, sums, soff, nb, st) != 0) return -1;
        hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, out, soff);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// out = exclusive scan of in (n entries); returns the total in *total_host
// (synchronous).  Uses ctx scratch slots [s0, s0 + 4).
int scan_exclusive_i64(Ctx* c, int s0, const int64_t* in, int64_t* out, int64_t n, int64_t* total_host,
                       hipStream_t st) {
    if (total_host) *total_host = 0;
    if (n <= 0) return 0;
    if (scan_rec(c, s0, in, out, n, st) != 0) return -1;
    if (total_host) {
        int64_t last_in = 0, last_out = 0;
        if (hipMemcpyAsync(&last_in, in + n - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&last_out, out + n - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("scan: total readback failed");
            return -1;
        }
        *total_host = last_in + last_out;
    }
    return 0;
}

// ------------------------------------------------------------- boxes and grid

struct Box {
    const double* lo;   // [n][3]
    const double* hi;   // [n][3]
};

__global__ void aabb_kernel(int n, const double* __restrict__ x, const int* __restrict__ conn,
                            const double* __restrict__ rad, double margin, double* __restrict__ lo,
                            double* __restrict__ hi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = conn[2 * i], b = conn[2 * i + 1];
    const double e = rad[i] + margin;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double p = x[3 * a + k], q = x[3 * b + k];
        // NaN propagates as through np.minimum / np.maximum (fmin and fmax would drop it)
        const bool nan = isnan(p) || isnan(q);
        lo[3 * i + k] = nan ? p + q : fmin(p, q) - e;
        hi[3 * i + k] = nan ? p + q : fmax(p, q) + e;
    }
}

// Per-workgroup reduction of the boxes (fixed-shape tree, no atomics, so the
// sums are deterministic): [0..2] min lo, [3..5] max hi, [6..8] max extent per
// axis, [9] sum over boxes of the largest extent, [10] boxes with a
// non-finite bound.  The host folds the partials, refuses non-finite boxes
// before any cell loop runs, and sizes the grid from the bounds.
constexpr int BRED = 11;
__global__ void __launch_bounds__(256)
box_reduce_kernel(int n, const double* __restrict__ lo, const double* __restrict__ hi, double* __restrict__ part) {
    __shared__ double sh[BRED][256];
    double v[BRED];
#pragma unroll
    for (int k = 0; k < 3; ++k) { v[k] = HUGE_VAL; v[3 + k] = -HUGE_VAL; v[6 + k] = 0.0; }
    v[9] = 0.0;
    v[10] = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double mx = 0.0;
        bool bad = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double l = lo[3 * i + k], h = hi[3 * i + k];
            bad |= !(isfinite(l) && isfinite(h));
            v[k] = fmin(v[k], l);
            v[3 + k] = fmax(v[3 + k], h);
            const double e = h - l;
            v[6 + k] = fmax(v[6 + k], e);
            mx = k == 0 ? e : fmax(mx, e);
        }
        v[9] += mx;
        v[10] += bad ? 1.0 : 0.0;
    }
#pragma unroll
    for (int q = 0; q < BRED; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const int t = threadIdx.x;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sh[k][t] = fmin(sh[k][t], sh[k][t + w]);
                sh[3 + k][t] = fmax(sh[3 + k][t], sh[3 + k][t + w]);
                sh[6 + k][t] = fmax(sh[6 + k][t], sh[6 + k][t + w]);
            }
            sh[9][t] += sh[9][t + w];
            sh[10][t] += sh[10][t + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < BRED) part[(size_t)blockIdx.x * BRED + threadIdx.x] = sh[threadIdx.x][0];
}

__device__ __forceinline__ int64_t cell_of(double v, double inv) { return (int64_t)floor(v * inv); }

struct Grid {
    double inv;
    long long o[3];   // min cell index per axis
    long long d[3];   // extent per axis
};

__device__ __forceinline__ int64_t cell_lin(const Grid& g, int64_t cx, int64_t cy, int64_t cz) {
    return ((cx - g.o[0]) * g.d[1] + (cy - g.o[1])) * g.d[2] + (cz - g.o[2]);
}

__global__ void grid_count_kernel(int n, Box B, Grid g, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t l[3], h[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { l[k] = cell_of(B.lo[3 * i + k], g.inv); h[k] = cell_of(B.hi[3 * i + k], g.inv); }
    for (int64_t cx = l[0]; cx <= h[0]; ++cx)
        for (int64_t cy = l[1]; cy <= h[1]; ++cy)
            for (int64_t cz = l[2]; cz <= h[2]; ++cz) atomicAdd(&cnt[cell_lin(g, cx, cy, cz)], 1);
}

__global__ void i32_to_i64_kernel(int64_t n, const int* __restrict__ a, int64_t* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void grid_fill_kernel(int n, Box B, Grid g, const int64_t* __restrict__ start, int* __restrict__ cursor,
                                 int* __restrict__ members) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t l[3], h[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { l[k] = cell_of(B.lo[3 * i + k], g.inv); h[k] = cell_of(B.hi[3 * i + k], g.inv); }
    for (int64_t cx = l[0]; cx <= h[0]; ++cx)
        for (int64_t cy = l[1]; cy <= h[1]; ++cy)
            for (int64_t cz = l[2]; cz <= h[2]; ++cz) {
                int64_t c = cell_lin(g, cx, cy, cz);
                int slot = atomicAdd(&cursor[c], 1);
                members[start[c] + slot] = i;
            }
}

struct Filt {
    const int* conn;      // shared-node filter when non-null
    const int* layer;     // same-layer filter when non-null
};

// One wavefront per grid cell: the cell's members are staged in LDS (up to
// CELL_LDS of them; larger cells read the boxes from global memory) and the
// wave's lanes stride over the cell's member pairs, testing each unordered
// pair once: inclusive overlap, then "this is the cell of max(lo_i, lo_j)",
// the one cell two overlapping boxes must share, then the filters.  Pass 0 counts the
// cell's pairs; pass 1 writes their keys (i << key_bits | j, i < j) at the
// cell's offset, compacted by ballot.  The member order inside a cell depends
// on the fill's atomics, so the keys are sorted afterwards.
constexpr int CELL_LDS = 128;

__global__ void __launch_bounds__(64)
cell_pairs_kernel(int64_t ncell, Box B, Grid g, const int64_t* __restrict__ start, const int* __restrict__ members,
                  Filt f, int pass, int key_bits, int64_t* __restrict__ count, const int64_t* __restrict__ off,
                  uint64_t* __restrict__ keys) {
    __shared__ double slo[3][CELL_LDS], shi[3][CELL_LDS];
    __shared__ int sid[CELL_LDS], sc0[CELL_LDS], sc1[CELL_LDS], slay[CELL_LDS];
    const int64_t c = blockIdx.x;
    if (c >= ncell) return;
    const int lane = threadIdx.x;
    const int64_t b0 = start[c];
    const int64_t m = start[c + 1] - b0;
    if (m < 2) {
        if (!pass && lane == 0) count[c] = 0;
        return;
    }
    const int64_t cz = g.o[2] + c % g.d[2];
    const int64_t cyx = c / g.d[2];
    const int64_t cy = g.o[1] + cyx % g.d[1];
    const int64_t cx = g.o[0] + cyx / g.d[1];
    const bool staged = m <= CELL_LDS;
    if (staged) {
        for (int q = lane; q < m; q += 64) {
            const int i = members[b0 + q];
            sid[q] = i;
#pragma unroll
            for (int k = 0; k < 3; ++k) { slo[k][q] = B.lo[3 * i + k]; shi[k][q] = B.hi[3 * i + k]; }
            sc0[q] = f.conn ? f.conn[2 * i] : 0;
            sc1[q] = f.conn ? f.conn[2 * i + 1] : 0;
            slay[q] = f.layer ? f.layer[i] : -1;
        }
    }
    __syncthreads();
    const int64_t base = pass ? off[c] : 0;
    int64_t running = 0;
    // one pair test: inclusive overlap, canonical cell, filters
    auto test = [&](int i, int j, const double* li, const double* hi_, const double* lj, const double* hj, int ci0,
                    int ci1, int cj0, int cj1, int lai, int laj) -> bool {
        (void)i; (void)j;
        bool keep = li[0] <= hj[0] && li[1] <= hj[1] && li[2] <= hj[2] && lj[0] <= hi_[0] && lj[1] <= hi_[1] &&
                    lj[2] <= hi_[2];
        keep = keep && cell_of(fmax(li[0], lj[0]), g.inv) == cx && cell_of(fmax(li[1], lj[1]), g.inv) == cy &&
               cell_of(fmax(li[2], lj[2]), g.inv) == cz;
        if (keep && f.conn) keep = !(ci0 == cj0 || ci0 == cj1 || ci1 == cj0 || ci1 == cj1);
        if (keep && f.layer) keep = !(lai == laj && lai >= 0);
        return keep;
    };
    auto emit = [&](bool keep, int i, int j) {
        const uint64_t mask = __ballot(keep);
        if (pass && keep) {
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            const uint64_t lo_id = (uint64_t)(i < j ? i : j), hi_id = (uint64_t)(i < j ? j : i);
            keys[base + running + rank] = (lo_id << key_bits) | hi_id;
        }
        running += __popcll(mask);
    };
    if (staged) {
        // the m (m - 1) / 2 member pairs a < b in row-major order, lanes over
        // the pair index p; row a starts at S(a) = a (2m - a - 1) / 2
        const int M = (int)m;
        const int P = M * (M - 1) / 2;
        const float t = (float)(2 * M - 1);
        for (int p0 = 0; p0 < P; p0 += 64) {
            const int p = p0 + lane;
            bool keep = false;
            int i = 0, j = 0;
            if (p < P) {
                int a = (int)((t - sqrtf(t * t - 8.0f * (float)p)) * 0.5f);
                a = a < 0 ? 0 : (a > M - 2 ? M - 2 : a);
                while (a > 0 && a * (2 * M - a - 1) / 2 > p) --a;
                while (a < M - 2 && (a + 1) * (2 * M - a - 2) / 2 <= p) ++a;
                const int b = p - a * (2 * M - a - 1) / 2 + a + 1;
                i = sid[a];
                j = sid[b];
                double li[3], hi_[3], lj[3], hj[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    li[k] = slo[k][a]; hi_[k] = shi[k][a];
                    lj[k] = slo[k][b]; hj[k] = shi[k][b];
                }
                keep = test(i, j, li, hi_, lj, hj, sc0[a], sc1[a], sc0[b], sc1[b], slay[a], slay[b]);
            }
            emit(keep, i, j);
        }
    } else {
        // a cell with more members than the LDS stage holds: every ordered
        // slot (a, b) of the m x m square, boxes from global memory
        const int64_t mm = m * m;
        for (int64_t q0 = 0; q0 < mm; q0 += 64) {
            const int64_t q = q0 + lane;
            bool keep = false;
            int i = 0, j = 0;
            if (q < mm) {
                const int64_t a = q / m, b = q - a * m;
                if (a < b) {
                    i = members[b0 + a];
                    j = members[b0 + b];
                    double li[3], hi_[3], lj[3], hj[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        li[k] = B.lo[3 * i + k]; hi_[k] = B.hi[3 * i + k];
                        lj[k] = B.lo[3 * j + k]; hj[k] = B.hi[3 * j + k];
                    }
                    keep = test(i, j, li, hi_, lj, hj, f.conn ? f.conn[2 * i] : 0, f.conn ? f.conn[2 * i + 1] : 0,
                                f.conn ? f.conn[2 * j] : 0, f.conn ? f.conn[2 * j + 1] : 0,
                                f.layer ? f.layer[i] : -1, f.layer ? f.layer[j] : -1);
                }
            }
            emit(keep, i, j);
        }
    }
    if (!pass && lane == 0) count[c] = running;
}

__global__ void split_keys_kernel(int64_t n, const uint64_t* __restrict__ keys, int key_bits, int* __restrict__ I,
                                  int* __restrict__ J) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint64_t k = keys[q];
    I[q] = (int)(k >> key_bits);
    J[q] = (int)(k & ((1ull << key_bits) - 1ull));
}

// ------------------------------------------------------- prescreen and merge

__global__ void prescreen_kernel(int64_t n, const int* __restrict__ I, const int* __restrict__ J,
                                 const double* __restrict__ x, const int* __restrict__ conn,
                                 const double* __restrict__ rad, double extra, int64_t* __restrict__ keep) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = I[q], j = J[q];
    const int n0 = conn[2 * i], n1 = conn[2 * i + 1], m0 = conn[2 * j], m1 = conn[2 * j + 1];
    double ma[3], mb[3], da[3], db[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ma[k] = 0.5 * (x[3 * n0 + k] + x[3 * n1 + k]);
        mb[k] = 0.5 * (x[3 * m0 + k] + x[3 * m1 + k]);
        da[k] = x[3 * n1 + k] - x[3 * n0 + k];
        db[k] = x[3 * m1 + k] - x[3 * m0 + k];
    }
    const double ha = 0.5 * sqrt((da[0] * da[0] + da[1] * da[1]) + da[2] * da[2]);
    const double hb = 0.5 * sqrt((db[0] * db[0] + db[1] * db[1]) + db[2] * db[2]);
    const double dx = ma[0] - mb[0], dy = ma[1] - mb[1], dz = ma[2] - mb[2];
    const double md = sqrt((dx * dx + dy * dy) + dz * dz);
    const double mind = fmax(0.0, (md - ha) - hb);
    keep[q] = mind <= (rad[i] + rad[j]) + extra ? 1 : 0;
}

__global__ void compact_kernel(int64_t n, const int64_t* __restrict__ keep, const int64_t* __restrict__ pos,
                               const int* __restrict__ I, const int* __restrict__ J, int* __restrict__ I2,
                               int* __restrict__ J2) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || !keep[q]) return;
    I2[pos[q]] = I[q];
    J2[pos[q]] = J[q];
}

__global__ void rsum_kernel(int64_t n, const int* __restrict__ I, const int* __restrict__ J,
                            const double* __restrict__ rad, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = rad[I[q]] + rad[J[q]];
}

__device__ __forceinline__ int64_t pkey(int a, int b) {
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((int64_t)lo << 32) | (int64_t)(uint32_t)hi;
}

// existing pair -> INACTIVE when its key is not among the candidates
__global__ void deactivate_kernel(int64_t np, const int* __restrict__ ea, const int* __restrict__ eb, int64_t nc,
                                  const int* __restrict__ I, const int* __restrict__ J, int8_t* __restrict__ status) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int64_t k = pkey(ea[p], eb[p]);
    int64_t lo = 0, hi = nc;
    while (lo < hi) {
        int64_t m = (lo + hi) >> 1;
        int64_t km = pkey(I[m], J[m]);
        if (km < k) lo = m + 1; else hi = m;
    }
    const bool found = lo < nc && pkey(I[lo], J[lo]) == k;
    if (!found) status[p] = ST_INACTIVE;
}

// candidate -> new (1) when its key is not among the sorted existing keys
__global__ void newflag_kernel(int64_t nc, const int* __restrict__ I, const int* __restrict__ J, int64_t nk,
                               const int64_t* __restrict__ keys, int64_t* __restrict__ isnew) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    const int64_t k = pkey(I[q], J[q]);
    int64_t lo = 0, hi = nk;
    while (lo < hi) {
        int64_t m = (lo + hi) >> 1;
        if (keys[m] < k) lo = m + 1; else hi = m;
    }
    isnew[q] = (lo < nk && keys[lo] == k) ? 0 : 1;
}

__global__ void append_kernel(int64_t nc, const int64_t* __restrict__ isnew, const int64_t* __restrict__ pos,
                              const int* __restrict__ I, const int* __restrict__ J, const int* __restrict__ conn,
                              const double* __restrict__ rad, const double* __restrict__ crad, PairSoA P,
                              int* __restrict__ ea, int* __restrict__ eb, int64_t base, int64_t* __restrict__ newkeys) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc || !isnew[q]) return;
    const int64_t p = base + pos[q];
    const int i = I[q], j = J[q];
    const int64_t cap = P.cap;
    ea[p] = i; eb[p] = j;
    P.na0[p] = conn[2 * i]; P.na1[p] = conn[2 * i + 1];
    P.nb0[p] = conn[2 * j]; P.nb1[p] = conn[2 * j + 1];
    P.ra[p] = rad[i]; P.rb[p] = rad[j];
    P.s[p] = 0.0; P.t[p] = 0.0; P.gap[p] = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) { P.n[k * cap + p] = 0.0; P.t1[k * cap + p] = 0.0; }
    P.p_n[p] = 0.0; P.k_pen[p] = 0.0; P.k_t[p] = 0.0; P.qtn[p] = 0.0; P.diss[p] = 0.0;
    P.z[p] = 0.0; P.z[cap + p] = 0.0;
    P.status[p] = ST_INACTIVE;
    P.stick[p] = 1;
    P.sinc[p] = -1;
    // coating: core radii of the elements, fresh state (_contact_pair.py:57-62)
    P.cra[p] = crad[i]; P.crb[p] = crad[j];
    P.cc[p] = 0.0; P.ccp[p] = 0.0;
    P.cz[p] = 0.0; P.cz[cap + p] = 0.0;
    P.cstick[p] = 1;
    P.cqtn[p] = 0.0; P.cdiss[p] = 0.0;
    P.cneed[p] = 0;
    newkeys[pos[q]] = pkey(i, j);
}

// merge two sorted key arrays (distinct keys) into out
__global__ void merge_kernel(int64_t na, const int64_t* __restrict__ a, int64_t nb, const int64_t* __restrict__ b,
                             int64_t* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= na + nb) return;
    const bool from_a = q < na;
    const int64_t k = from_a ? a[q] : b[q - na];
    const int64_t* o = from_a ? b : a;
    const int64_t no = from_a ? nb : na;
    int64_t lo = 0, hi = no;
    while (lo < hi) {
        int64_t m = (lo + hi) >> 1;
        if (o[m] < k) lo = m + 1; else hi = m;
    }
    out[(from_a ? q : q - na) + lo] = k;
}

__global__ void keys_of_pairs_kernel(int64_t np, const int* __restrict__ ea, const int* __restrict__ eb,
                                     int64_t* __restrict__ keys) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) keys[p] = pkey(ea[p], eb[p]);
}

static unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// scratch slots: 0..5 scan recursion; 6 boxes; 7 box-reduction partials; 8 grid counts/start;
// 9 members; 10 per-cell pair counts/offsets; 11 raw I/J; 12 filtered I/J; 13 flags/positions;
// 14 sorted existing keys; 15 new keys / merge output; 16 pair keys (emitted, sorted);
// 17 radix-sort scratch
enum { S_BOX = 6, S_EXT = 7, S_GRID = 8, S_MEM = 9, S_CNT = 10, S_RAW = 11, S_FILT = 12, S_FLAG = 13,
       S_KEYS = 14, S_NEWK = 15, S_PKEY = 16, S_SORT = 17 };

#define BP_TRY(expr)                                                                 \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return -1;                                                               \
        }                                                                            \
    } while (0)

static hipError_t sort_keys_u64(Ctx* c, const uint64_t* in, uint64_t* out, int64_t n, int end_bit,
                                hipStream_t st) {
    if (n <= 0) return hipSuccess;
    size_t tmp = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, tmp, in, out, (size_t)n, 0, end_bit, st);
    if (e != hipSuccess) return e;
    if ((e = c->bp[S_SORT].ensure(std::max<size_t>(tmp, 1))) != hipSuccess) return e;
    return rocprim::radix_sort_keys(c->bp[S_SORT].p, tmp, in, out, (size_t)n, 0, end_bit, st);
}

// raw (filter=0), shared-node/layer filtered (filter=1) or shared-node
// filtered (filter=2) broadphase of all elements; result I/J (n_out, sorted by
// (i, j)) left in slot S_RAW.
static int run_broadphase(Ctx* c, const double* coords, double margin, double cell_size, int filter,
                          const int* layer, int64_t* n_out, hipStream_t st) {
    const int n = c->n_elems;
    *n_out = 0;
    if (n < 2) return 0;
    BP_TRY(c->bp[S_BOX].ensure(sizeof(double) * 6 * (size_t)n));
    double* lo = (double*)c->bp[S_BOX].p;
    double* hi = lo + 3 * (size_t)n;
    const unsigned nbs = std::min<unsigned>(nblk(n), 1024);
    BP_TRY(c->bp[S_EXT].ensure(sizeof(double) * BRED * (size_t)nbs));
    double* part = (double*)c->bp[S_EXT].p;
    hipLaunchKernelGGL(aabb_kernel, dim3(nblk(n)), dim3(256), 0, st, n, coords, (const int*)c->d_conn.p,
                       (const double*)c->d_rad.p, margin, lo, hi);
    BP_TRY(hipGetLastError());
    hipLaunchKernelGGL(box_reduce_kernel, dim3(nbs), dim3(256), 0, st, n, lo, hi, part);
    BP_TRY(hipGetLastError());
    std::vector<double> ph((size_t)nbs * BRED);
    BP_TRY(hipMemcpyAsync(ph.data(), part, sizeof(double) * ph.size(), hipMemcpyDeviceToHost, st));
    BP_TRY(hipStreamSynchronize(st));
    double bmin[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, bmax[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    double emax[3] = {0.0, 0.0, 0.0}, esum = 0.0, nbad = 0.0;
    for (unsigned b = 0; b < nbs; ++b) {
        const double* v = ph.data() + (size_t)b * BRED;
        for (int k = 0; k < 3; ++k) {
            bmin[k] = std::min(bmin[k], v[k]);
            bmax[k] = std::max(bmax[k], v[3 + k]);
            emax[k] = std::max(emax[k], v[6 + k]);
        }
        esum += v[9];
        nbad += v[10];
    }
    if (nbad > 0.0 || !std::isfinite(esum)) {
        set_error("broadphase: non-finite coordinates in " + std::to_string((long long)nbad) +
                  " element boxes (the displacement has diverged)");
        return -1;
    }
    if (!(cell_size > 0.0)) cell_size = std::max(esum / n * c->cell_factor, 1e-30);   // tuned default
    // enlarge the cell (the set does not depend on it) until the grid and the
    // per-box cell ranges stay bounded; all cell indices are formed in double
    // and checked to be exact integers before they become grid extents
    Grid g{};
    double cells = 0.0;
    for (int attempt = 0;; ++attempt) {
        if (attempt == 60) {
            set_error("broadphase: could not size the grid");
            return -1;
        }
        g.inv = 1.0 / cell_size;
        cells = 1.0;
        double per_box = 1.0;
        bool exact = true;
        for (int k = 0; k < 3; ++k) {
            const double a = std::floor(bmin[k] * g.inv), b = std::floor(bmax[k] * g.inv);
            exact &= std::fabs(a) < 9.0e15 && std::fabs(b) < 9.0e15;
            g.o[k] = exact ? (long long)a : 0;
            g.d[k] = exact ? (long long)(b - a) + 1 : 0;
            cells *= b - a + 1.0;
            per_box *= std::floor(emax[k] * g.inv) + 2.0;
        }
        if (exact && cells <= (double)(1LL << 26) && per_box <= 4096.0 && cells <= 64.0 * n + (1 << 20)) break;
        cell_size *= 2.0;
    }
    const int64_t ncell = g.d[0] * g.d[1] * g.d[2];
    BP_TRY(c->bp[S_GRID].ensure(sizeof(int64_t) * (size_t)(2 * ncell + 2) + sizeof(int) * (size_t)(ncell + 1)));
    int64_t* gcnt64 = (int64_t*)c->bp[S_GRID].p;
    int64_t* gstart = gcnt64 + ncell + 1;
    int* gcnt = (int*)(gstart + ncell + 1);
    BP_TRY(launch_fill_i32((size_t)ncell, gcnt, 0, st));
    Box B{lo, hi};
    hipLaunchKernelGGL(grid_count_kernel, dim3(nblk(n)), dim3(256), 0, st, n, B, g, gcnt);
    hipLaunchKernelGGL(i32_to_i64_kernel, dim3(nblk(ncell)), dim3(256), 0, st, ncell, gcnt, gcnt64);
    BP_TRY(hipGetLastError());
    int64_t total = 0;
    if (scan_exclusive_i64(c, 0, gcnt64, gstart, ncell, &total, st) != 0) return -1;
    BP_TRY(hipMemcpyAsync(gstart + ncell, &total, 8, hipMemcpyHostToDevice, st));
    BP_TRY(c->bp[S_MEM].ensure(sizeof(int) * (size_t)std::max<int64_t>(total, 1)));
    int* members = (int*)c->bp[S_MEM].p;
    BP_TRY(launch_fill_i32((size_t)ncell, gcnt, 0, st));
    hipLaunchKernelGGL(grid_fill_kernel, dim3(nblk(n)), dim3(256), 0, st, n, B, g, gstart, gcnt, members);
    BP_TRY(hipGetLastError());

    // per-cell pair counts, their offsets, the keys, sorted
    BP_TRY(c->bp[S_CNT].ensure(sizeof(int64_t) * 2 * (size_t)(ncell + 1)));
    int64_t* cnt = (int64_t*)c->bp[S_CNT].p;
    int64_t* off = cnt + ncell + 1;
    Filt f{filter ? (const int*)c->d_conn.p : nullptr, filter == 1 ? layer : nullptr};
    int key_bits = 1;
    while ((1LL << key_bits) < (int64_t)n) ++key_bits;
    hipLaunchKernelGGL(cell_pairs_kernel, dim3((unsigned)ncell), dim3(64), 0, st, ncell, B, g, gstart, members, f, 0,
                       key_bits, cnt, off, (uint64_t*)nullptr);
    BP_TRY(hipGetLastError());
    int64_t npairs = 0;
    if (scan_exclusive_i64(c, 0, cnt, off, ncell, &npairs, st) != 0) return -1;
    BP_TRY(c->bp[S_PKEY].ensure(sizeof(uint64_t) * 2 * (size_t)std::max<int64_t>(npairs, 1)));
    uint64_t* keys = (uint64_t*)c->bp[S_PKEY].p;
    uint64_t* sorted = keys + std::max<int64_t>(npairs, 1);
    BP_TRY(c->bp[S_RAW].ensure(sizeof(int) * 2 * (size_t)std::max<int64_t>(npairs, 1)));
    int* I = (int*)c->bp[S_RAW].p;
    int* J = I + std::max<int64_t>(npairs, 1);
    if (npairs > 0) {
        hipLaunchKernelGGL(cell_pairs_kernel, dim3((unsigned)ncell), dim3(64), 0, st, ncell, B, g, gstart, members, f,
                           1, key_bits, cnt, off, keys);
        BP_TRY(hipGetLastError());
        BP_TRY(sort_keys_u64(c, keys, sorted, npairs, 2 * key_bits, st));
        hipLaunchKernelGGL(split_keys_kernel, dim3(nblk(npairs)), dim3(256), 0, st, npairs, (const uint64_t*)sorted,
                           key_bits, I, J);
        BP_TRY(hipGetLastError());
    }
    c->raw_n = npairs;
    c->raw_cap = std::max<int64_t>(npairs, 1);
    c->last_cell = cell_size;
    *n_out = npairs;
    return 0;
}

// The pair-list update of DetectCandidatesProcess (contact/_manager_process.py:219-250)
// for the sorted candidate list I2/J2 (n2): existing pairs outside it become
// INACTIVE (state kept), candidates not yet in the list are appended with a
// fresh state.  Synchronous.
static int merge_candidates(Ctx* c, const int* I2, const int* J2, int64_t n2, hipStream_t st) {

    // existing pairs outside the candidates -> INACTIVE
    const int64_t np = c->n_pairs;
    if (np > 0) {
        hipLaunchKernelGGL(deactivate_kernel, dim3(nblk(np)), dim3(256), 0, st, np, c->ea, c->eb, n2, I2, J2,
                           c->P.status);
        BP_TRY(hipGetLastError());
    }
    // sorted keys of the existing pairs
    if (!c->keys_valid) {
        BP_TRY(c->bp[S_KEYS].ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(np, 1)));
        if (np > 0) {
            std::vector<int64_t> hk(np);
            std::vector<int> a(np), b(np);
            BP_TRY(hipMemcpy(a.data(), c->ea, 4 * np, hipMemcpyDeviceToHost));
            BP_TRY(hipMemcpy(b.data(), c->eb, 4 * np, hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < np; ++p) {
                int lo_ = std::min(a[p], b[p]), hi_ = std::max(a[p], b[p]);
                hk[p] = ((int64_t)lo_ << 32) | (int64_t)(uint32_t)hi_;
            }
            std::sort(hk.begin(), hk.end());
            BP_TRY(hipMemcpy(c->bp[S_KEYS].p, hk.data(), 8 * np, hipMemcpyHostToDevice));
        }
        c->keys_valid = true;
    }
    int64_t n_new = 0;
    if (n2 > 0) {
        BP_TRY(c->bp[S_NEWK].ensure(sizeof(int64_t) * (3 * (size_t)n2 + (size_t)np + 2)));
        int64_t* isnew = (int64_t*)c->bp[S_NEWK].p;
        int64_t* pos = isnew + n2;
        int64_t* newkeys = pos + n2;
        hipLaunchKernelGGL(newflag_kernel, dim3(nblk(n2)), dim3(256), 0, st, n2, I2, J2, np,
                           (const int64_t*)c->bp[S_KEYS].p, isnew);
        BP_TRY(hipGetLastError());
        if (scan_exclusive_i64(c, 0, isnew, pos, n2, &n_new, st) != 0) return -1;
        if (n_new > 0) {
            if (c->grow_pairs(np + n_new) != 0) return -1;
            hipLaunchKernelGGL(append_kernel, dim3(nblk(n2)), dim3(256), 0, st, n2, isnew, pos, I2, J2,
                               (const int*)c->d_conn.p, (const double*)c->d_rad.p, (const double*)c->d_crad.p, c->P,
                               c->ea, c->eb, np, newkeys);
            BP_TRY(hipGetLastError());
            // merged sorted keys
            DBuf merged;
            BP_TRY(merged.ensure(sizeof(int64_t) * (size_t)(np + n_new)));
            hipLaunchKernelGGL(merge_kernel, dim3(nblk(np + n_new)), dim3(256), 0, st, np,
                               (const int64_t*)c->bp[S_KEYS].p, n_new, (const int64_t*)newkeys,
                               (int64_t*)merged.p);
            BP_TRY(hipGetLastError());
            BP_TRY(hipStreamSynchronize(st));
            c->bp[S_KEYS].release();
            c->bp[S_KEYS] = merged;
            c->n_pairs = np + n_new;
        }
    }
    BP_TRY(hipStreamSynchronize(st));
    c->pattern_valid = false;
    c->lp_on = false;
    return 0;
}

}  // namespace xk

using namespace xk;

struct xk_ctx : public Ctx {};

extern "C" {

int xk_broadphase(xk_ctx* c, const double* coords, double margin, double cell_size, int32_t* pairs_out, int64_t cap,
                  int64_t* n_out, void* stream) {
    if (!c || !coords || !n_out) {
        set_error("xk_broadphase: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    int64_t n = 0;
    if (run_broadphase(c, coords, margin, cell_size, 0, nullptr, &n, st) != 0) return -1;
    *n_out = n;
    if (pairs_out && n > 0) {
        if (cap < n) {
            set_error("xk_broadphase: output capacity too small");
            return -1;
        }
        const int* I = (const int*)c->bp[S_RAW].p;
        const int* J = I + c->raw_cap;
        BP_TRY(hipMemcpyAsync(pairs_out, I, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
        BP_TRY(hipMemcpyAsync(pairs_out + n, J, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
        BP_TRY(hipStreamSynchronize(st));
    }
    return 0;
}

int xk_prescreen_radii(xk_ctx* c, const double* coords, double margin, double cell_size, const int32_t* elem_layer,
                       double* rsum_host, int64_t cap, int64_t* n_out, void* stream) {
    if (!c || !coords || !n_out) {
        set_error("xk_prescreen_radii: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    int64_t n = 0;
    if (run_broadphase(c, coords, margin, cell_size, 1, elem_layer, &n, st) != 0) return -1;
    c->filtered_valid = true;
    *n_out = n;
    if (rsum_host && n > 0) {
        if (cap < n) {
            set_error("xk_prescreen_radii: output capacity too small");
            return -1;
        }
        BP_TRY(c->bp[S_FLAG].ensure(sizeof(double) * (size_t)n));
        const int* I = (const int*)c->bp[S_RAW].p;
        const int* J = I + c->raw_cap;
        hipLaunchKernelGGL(rsum_kernel, dim3(nblk(n)), dim3(256), 0, st, n, I, J, (const double*)c->d_rad.p,
                           (double*)c->bp[S_FLAG].p);
        BP_TRY(hipMemcpyAsync(rsum_host, c->bp[S_FLAG].p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st));
        BP_TRY(hipStreamSynchronize(st));
    }
    return 0;
}

int xk_detect_candidates(xk_ctx* c, const double* coords, double margin, double cell_size, const int32_t* elem_layer,
                         double prescreen_extra, int64_t* n_raw, int64_t* n_candidates, void* stream) {
    if (c && c->lp_sync((hipStream_t)stream, false) != 0) return -1;
    if (!c || !coords) {
        set_error("xk_detect_candidates: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    int64_t n1 = 0;
    if (c->filtered_valid) {
        n1 = c->raw_n;  // reuse the filtered list of the preceding xk_prescreen_radii
    } else if (run_broadphase(c, coords, margin, cell_size, 1, elem_layer, &n1, st) != 0) {
        return -1;
    }
    c->filtered_valid = false;
    if (n_raw) *n_raw = n1;
    const int* I = (const int*)c->bp[S_RAW].p;
    const int* J = I + c->raw_cap;
    // midpoint prescreen (stable compaction)
    int64_t n2 = n1;
    const int* I2 = I;
    const int* J2 = J;
    if (prescreen_extra >= 0.0 && n1 > 0) {
        BP_TRY(c->bp[S_FLAG].ensure(sizeof(int64_t) * 2 * (size_t)n1));
        int64_t* keep = (int64_t*)c->bp[S_FLAG].p;
        int64_t* pos = keep + n1;
        hipLaunchKernelGGL(prescreen_kernel, dim3(nblk(n1)), dim3(256), 0, st, n1, I, J, coords,
                           (const int*)c->d_conn.p, (const double*)c->d_rad.p, prescreen_extra, keep);
        BP_TRY(hipGetLastError());
        if (scan_exclusive_i64(c, 0, keep, pos, n1, &n2, st) != 0) return -1;
        BP_TRY(c->bp[S_FILT].ensure(sizeof(int) * 2 * (size_t)std::max<int64_t>(n2, 1)));
        int* If = (int*)c->bp[S_FILT].p;
        int* Jf = If + std::max<int64_t>(n2, 1);
        hipLaunchKernelGGL(compact_kernel, dim3(nblk(n1)), dim3(256), 0, st, n1, keep, pos, I, J, If, Jf);
        BP_TRY(hipGetLastError());
        I2 = If;
        J2 = Jf;
        c->cand_I = If;
        c->cand_J = Jf;
    } else {
        c->cand_I = I;
        c->cand_J = J;
    }
    c->last_cand = n2;
    if (n_candidates) *n_candidates = n2;
    return merge_candidates(c, I2, J2, n2, st);
}

int xk_strategy_detect(xk_ctx* c, const double* coords, double margin, double cell_size, int shared_node_filter,
                       int64_t* n_pairs, void* stream) {
    if (!c || !coords || !n_pairs) {
        set_error("xk_strategy_detect: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (c->lp_sync(st, false) != 0) return -1;
    BP_TRY(hipStreamSynchronize(st));
    // a new list: every candidate is a new pair with a fresh state
    c->n_pairs = 0;
    c->keys_valid = false;
    c->filtered_valid = false;
    c->lp_on = false;
    int64_t n = 0;
    if (run_broadphase(c, coords, margin, cell_size, shared_node_filter ? 2 : 0, nullptr, &n, st) != 0) return -1;
    const int* I = (const int*)c->bp[S_RAW].p;
    const int* J = I + c->raw_cap;
    c->cand_I = I;
    c->cand_J = J;
    c->last_cand = n;
    if (merge_candidates(c, I, J, n, st) != 0) return -1;
    *n_pairs = c->n_pairs;
    return 0;
}

int xk_last_candidates(xk_ctx* c, int32_t* pairs_out_host, int64_t cap) {
    if (!c || !pairs_out_host) return -1;
    const int64_t n = c->last_cand;
    if (cap < n) {
        set_error("xk_last_candidates: capacity too small");
        return -1;
    }
    if (n == 0) return 0;
    BP_TRY(hipDeviceSynchronize());
    BP_TRY(hipMemcpy(pairs_out_host, c->cand_I, 4 * (size_t)n, hipMemcpyDeviceToHost));
    BP_TRY(hipMemcpy(pairs_out_host + n, c->cand_J, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
