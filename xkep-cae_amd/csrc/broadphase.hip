// GPU broadphase and candidate management (DetectCandidatesProcess).
//
// Output contract: the exact set of element pairs (i < j) whose FP64
// axis-aligned boxes
//     lo = min(x0, x1) - (r + margin),  hi = max(x0, x1) + (r + margin)
// overlap inclusively on all three axes -- the set the reference's
// _broadphase_aabb returns (xkep_cae/contact/_broadphase.py:13-96), here in
// sorted (i, j) order.
//
// Method: a dense uniform grid.  Every box is binned into the cells its
// floor(lo * inv) .. floor(hi * inv) range covers (counting sort by cell).
// Lane i then visits the members j > i of its own cells and keeps a pair
// only in the cell holding floor(max(lo_i, lo_j) * inv), the one cell both
// boxes are guaranteed to share when they overlap -- so every pair is found
// exactly once and the cell size only changes the work, never the set.  Each
// lane writes its partners to a segment placed by an exclusive scan of the
// counts and sorts it, which makes the list canonical.
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
    if (slot >= 16) {
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
                            double* __restrict__ hi, double* __restrict__ ext) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = conn[2 * i], b = conn[2 * i + 1];
    const double e = rad[i] + margin;
    double mx = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double p = x[3 * a + k], q = x[3 * b + k];
        double l = fmin(p, q) - e, h = fmax(p, q) + e;
        lo[3 * i + k] = l;
        hi[3 * i + k] = h;
        mx = k == 0 ? (h - l) : fmax(mx, h - l);
    }
    ext[i] = mx;
}

// per-block partial sums of ext (fixed-shape tree: deterministic)
__global__ void __launch_bounds__(256) sum_kernel(int n, const double* __restrict__ v, double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__device__ __forceinline__ int64_t cell_of(double v, double inv) { return (int64_t)floor(v * inv); }

// bounds of cell indices over all boxes; also flags too-wide boxes
__global__ void cell_bounds_kernel(int n, const double* __restrict__ lo, const double* __restrict__ hi, double inv,
                                   long long* __restrict__ bmin, long long* __restrict__ bmax,
                                   long long* __restrict__ maxcount) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long cnt = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        long long l = cell_of(lo[3 * i + k], inv), h = cell_of(hi[3 * i + k], inv);
        atomicMin(&bmin[k], l);
        atomicMax(&bmax[k], h);
        cnt *= (h - l + 1);
    }
    atomicMax(maxcount, cnt);
}

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

// pass 0: count partners; pass 1: write them and sort the segment
__global__ void pairs_kernel(int n, Box B, Grid g, const int64_t* __restrict__ start, const int* __restrict__ members,
                             Filt f, int pass, int64_t* __restrict__ count, const int64_t* __restrict__ off,
                             int* __restrict__ outI, int* __restrict__ outJ) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double li[3], hi_[3];
    int64_t l[3], h[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        li[k] = B.lo[3 * i + k];
        hi_[k] = B.hi[3 * i + k];
        l[k] = cell_of(li[k], g.inv);
        h[k] = cell_of(hi_[k], g.inv);
    }
    int ci0 = 0, ci1 = 0, lay = -1;
    if (f.conn) { ci0 = f.conn[2 * i]; ci1 = f.conn[2 * i + 1]; }
    if (f.layer) lay = f.layer[i];
    int64_t k_out = 0;
    const int64_t base = pass ? off[i] : 0;
    for (int64_t cx = l[0]; cx <= h[0]; ++cx)
        for (int64_t cy = l[1]; cy <= h[1]; ++cy)
            for (int64_t cz = l[2]; cz <= h[2]; ++cz) {
                const int64_t c = cell_lin(g, cx, cy, cz);
                for (int64_t q = start[c]; q < start[c + 1]; ++q) {
                    const int j = members[q];
                    if (j <= i) continue;
                    const double* lj = B.lo + 3 * j;
                    const double* hj = B.hi + 3 * j;
                    if (!(li[0] <= hj[0] && li[1] <= hj[1] && li[2] <= hj[2] && lj[0] <= hi_[0] &&
                          lj[1] <= hi_[1] && lj[2] <= hi_[2]))
                        continue;
                    // canonical cell: the cell of max(lo_i, lo_j)
                    if (cell_of(fmax(li[0], lj[0]), g.inv) != cx || cell_of(fmax(li[1], lj[1]), g.inv) != cy ||
                        cell_of(fmax(li[2], lj[2]), g.inv) != cz)
                        continue;
                    if (f.conn) {
                        const int cj0 = f.conn[2 * j], cj1 = f.conn[2 * j + 1];
                        if (ci0 == cj0 || ci0 == cj1 || ci1 == cj0 || ci1 == cj1) continue;
                    }
                    if (f.layer) {
                        const int lj_ = f.layer[j];
                        if (lay == lj_ && lay >= 0) continue;
                    }
                    if (pass) {
                        outI[base + k_out] = i;
                        outJ[base + k_out] = j;
                    }
                    ++k_out;
                }
            }
    if (!pass) {
        count[i] = k_out;
        return;
    }
    // insertion sort of this lane's partners by j
    for (int64_t a = 1; a < k_out; ++a) {
        int v = outJ[base + a];
        int64_t b = a - 1;
        while (b >= 0 && outJ[base + b] > v) {
            outJ[base + b + 1] = outJ[base + b];
            --b;
        }
        outJ[base + b + 1] = v;
    }
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

// scratch slots: 0..5 scan recursion; 6 boxes; 7 ext/partials; 8 grid counts/start;
// 9 members; 10 per-element counts/offsets; 11 raw I/J; 12 filtered I/J; 13 flags/positions;
// 14 sorted existing keys; 15 new keys / merge output
enum { S_BOX = 6, S_EXT = 7, S_GRID = 8, S_MEM = 9, S_CNT = 10, S_RAW = 11, S_FILT = 12, S_FLAG = 13,
       S_KEYS = 14, S_NEWK = 15 };

#define BP_TRY(expr)                                                                 \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return -1;                                                               \
        }                                                                            \
    } while (0)

// raw (filter=0) or shared-node/layer filtered (filter=1) broadphase of all
// elements; result I/J (n_out) left in slot S_RAW.
static int run_broadphase(Ctx* c, const double* coords, double margin, double cell_size, int filter,
                          const int* layer, int64_t* n_out, hipStream_t st) {
    const int n = c->n_elems;
    *n_out = 0;
    if (n < 2) return 0;
    BP_TRY(c->bp[S_BOX].ensure(sizeof(double) * 6 * (size_t)n));
    double* lo = (double*)c->bp[S_BOX].p;
    double* hi = lo + 3 * (size_t)n;
    const unsigned nbs = std::min<unsigned>(nblk(n), 1024);
    BP_TRY(c->bp[S_EXT].ensure(sizeof(double) * ((size_t)n + nbs + 8) + 64));
    double* ext = (double*)c->bp[S_EXT].p;
    double* part = ext + n;
    hipLaunchKernelGGL(aabb_kernel, dim3(nblk(n)), dim3(256), 0, st, n, coords, (const int*)c->d_conn.p,
                       (const double*)c->d_rad.p, margin, lo, hi, ext);
    BP_TRY(hipGetLastError());
    if (!(cell_size > 0.0)) {
        // tuned default: 1.5 x mean of the boxes' largest extent
        hipLaunchKernelGGL(sum_kernel, dim3(nbs), dim3(256), 0, st, n, ext, part);
        std::vector<double> ph(nbs);
        BP_TRY(hipMemcpyAsync(ph.data(), part, sizeof(double) * nbs, hipMemcpyDeviceToHost, st));
        BP_TRY(hipStreamSynchronize(st));
        double s = 0.0;
        for (double v : ph) s += v;
        cell_size = std::max(s / n * c->cell_factor, 1e-30);
    }
    Grid g{};
    long long* bnd = (long long*)(part + nbs + 1);
    for (int attempt = 0; attempt < 40; ++attempt) {
        g.inv = 1.0 / cell_size;
        long long init[7] = {LLONG_MAX, LLONG_MAX, LLONG_MAX, LLONG_MIN, LLONG_MIN, LLONG_MIN, 0};
        BP_TRY(hipMemcpyAsync(bnd, init, sizeof(init), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(cell_bounds_kernel, dim3(nblk(n)), dim3(256), 0, st, n, lo, hi, g.inv, bnd, bnd + 3,
                           bnd + 6);
        long long got[7];
        BP_TRY(hipMemcpyAsync(got, bnd, sizeof(got), hipMemcpyDeviceToHost, st));
        BP_TRY(hipStreamSynchronize(st));
        double cells = 1.0;
        for (int k = 0; k < 3; ++k) {
            g.o[k] = got[k];
            g.d[k] = got[3 + k] - got[k] + 1;
            cells *= (double)g.d[k];
        }
        // enlarge the cell (the set does not depend on it) until the grid and
        // the per-box cell ranges stay bounded
        if (cells <= (double)(1LL << 26) && got[6] <= 4096 && cells * 1.0 <= 64.0 * n + (1 << 20)) break;
        cell_size *= 2.0;
        if (attempt == 39) {
            set_error("broadphase: could not size the grid");
            return -1;
        }
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

    BP_TRY(c->bp[S_CNT].ensure(sizeof(int64_t) * 2 * (size_t)(n + 1)));
    int64_t* cnt = (int64_t*)c->bp[S_CNT].p;
    int64_t* off = cnt + n + 1;
    Filt f{filter ? (const int*)c->d_conn.p : nullptr, filter ? layer : nullptr};
    hipLaunchKernelGGL(pairs_kernel, dim3(nblk(n)), dim3(256), 0, st, n, B, g, gstart, members, f, 0, cnt, off,
                       (int*)nullptr, (int*)nullptr);
    BP_TRY(hipGetLastError());
    int64_t npairs = 0;
    if (scan_exclusive_i64(c, 0, cnt, off, n, &npairs, st) != 0) return -1;
    BP_TRY(c->bp[S_RAW].ensure(sizeof(int) * 2 * (size_t)std::max<int64_t>(npairs, 1)));
    int* I = (int*)c->bp[S_RAW].p;
    int* J = I + std::max<int64_t>(npairs, 1);
    hipLaunchKernelGGL(pairs_kernel, dim3(nblk(n)), dim3(256), 0, st, n, B, g, gstart, members, f, 1, cnt, off, I, J);
    BP_TRY(hipGetLastError());
    c->raw_n = npairs;
    c->raw_cap = std::max<int64_t>(npairs, 1);
    c->last_cell = cell_size;
    *n_out = npairs;
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
