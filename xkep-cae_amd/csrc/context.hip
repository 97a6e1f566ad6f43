// Host side of the gfx950 assembly library: device context, pair list,
// per-increment tangent pattern and the extern "C" entry points declared in
// include/xkamd.h.

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xk_common.h"
#include "../../include/xkamd.h"
#include "xk_internal.h"

namespace xk {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

#define XK_TRY(expr)                                                                 \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return -1;                                                               \
        }                                                                            \
    } while (0)

hipError_t DBuf::ensure(size_t nbytes) {
    if (nbytes <= bytes && p) return hipSuccess;
    if (p) {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return e;
        p = nullptr;
        bytes = 0;
    }
    size_t want = nbytes < 256 ? 256 : nbytes;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
}

void DBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

template <class T>
static hipError_t upload(DBuf& b, const std::vector<T>& v) {
    hipError_t e = b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// counting sort of (key, code) pairs into CSR lists, preserving insertion order
static void bucket(int n_keys, const std::vector<int>& keys, const std::vector<int>& codes,
                   std::vector<int>& ptr, std::vector<int>& src) {
    ptr.assign(n_keys + 1, 0);
    for (int k : keys) ptr[k + 1]++;
    for (int i = 0; i < n_keys; ++i) ptr[i + 1] += ptr[i];
    src.resize(keys.size());
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t q = 0; q < keys.size(); ++q) src[fill[keys[q]]++] = codes[q];
}

// bytes per pair of each pair-state array, and its component count
static const size_t PB_SZ[PB_COUNT] = {4, 4, 4, 4, 8, 8, 8, 8, 8, 24, 24, 8, 8, 8, 8, 8, 16, 1, 1, 4, 4, 4,
                                       8, 8, 8, 8, 16, 1, 8, 8, 1};
static int pb_comps(int k) { return (k == PB_N || k == PB_T1) ? 3 : ((k == PB_Z || k == PB_CZ) ? 2 : 1); }
// arrays the pair kernel writes (the others are the pair's identity)
static bool pb_written(int k) {
    return !(k == PB_NA0 || k == PB_NA1 || k == PB_NB0 || k == PB_NB1 || k == PB_RA || k == PB_RB ||
             k == PB_SINC || k == PB_EA || k == PB_EB || k == PB_CRA || k == PB_CRB);
}

static void bind_coating(PairSoA& P, std::vector<DBuf>& b) {
    P.cra = (double*)b[PB_CRA].p; P.crb = (double*)b[PB_CRB].p;
    P.cc = (double*)b[PB_CC].p; P.ccp = (double*)b[PB_CCP].p;
    P.cz = (double*)b[PB_CZ].p;
    P.cstick = (int8_t*)b[PB_CSTICK].p;
    P.cqtn = (double*)b[PB_CQTN].p; P.cdiss = (double*)b[PB_CDISS].p;
    P.cneed = (int8_t*)b[PB_CNEED].p;
}

static_assert(PB_COUNT <= PERM_SET_MAX, "one permute_set launch moves every pair-state array");
static_assert(PB_CRA + 9 == PB_COUNT, "the coating arrays are the last nine");

static PairSoA soa_of(std::vector<DBuf>& b, int64_t cap) {
    PairSoA P{};
    P.na0 = (int*)b[PB_NA0].p; P.na1 = (int*)b[PB_NA1].p;
    P.nb0 = (int*)b[PB_NB0].p; P.nb1 = (int*)b[PB_NB1].p;
    P.ra = (double*)b[PB_RA].p; P.rb = (double*)b[PB_RB].p;
    P.s = (double*)b[PB_S].p; P.t = (double*)b[PB_T].p; P.gap = (double*)b[PB_GAP].p;
    P.n = (double*)b[PB_N].p; P.t1 = (double*)b[PB_T1].p;
    P.p_n = (double*)b[PB_PN].p; P.k_pen = (double*)b[PB_KPEN].p; P.k_t = (double*)b[PB_KT].p;
    P.qtn = (double*)b[PB_QTN].p; P.diss = (double*)b[PB_DISS].p; P.z = (double*)b[PB_Z].p;
    P.status = (int8_t*)b[PB_STATUS].p; P.stick = (int8_t*)b[PB_STICK].p;
    P.sinc = (int*)b[PB_SINC].p;
    bind_coating(P, b);
    P.cap = cap;
    return P;
}

// The watch list and the coating are never on together (xk_lp_select refuses
// the coating, xk_ctx_set_coating ends the watch list first), so the coating
// arrays are neither compacted nor scattered back.
static bool pb_coating(int k) { return k >= PB_CRA; }

int Ctx::lp_compact(hipStream_t st) {
    wb.resize(PB_COUNT);
    const int64_t n = n_watch;
    PermuteSet set;
    for (int k = 0; k < PB_COUNT; ++k) {
        XK_TRY(wb[k].ensure(PB_SZ[k] * (size_t)std::max<int64_t>(n, 1)));
        if (pb_coating(k)) continue;
        const int comps = pb_comps(k);
        set.a[set.count] = pb[k].p;
        set.b[set.count] = wb[k].p;
        set.esz[set.count] = (int)(PB_SZ[k] / comps);
        set.comps[set.count] = comps;
        ++set.count;
    }
    XK_TRY(launch_permute_set(set, n, cap, n, (const int*)lp_watch.p, 0, st));
    PW = soa_of(wb, std::max<int64_t>(n, 1));
    return 0;
}

int Ctx::lp_sync(hipStream_t st, bool keep_on) {
    if (!lp_on) return 0;
    PermuteSet set;
    for (int k = 0; k < PB_COUNT; ++k) {
        if (!pb_written(k) || pb_coating(k)) continue;
        const int comps = pb_comps(k);
        set.a[set.count] = wb[k].p;
        set.b[set.count] = pb[k].p;
        set.esz[set.count] = (int)(PB_SZ[k] / comps);
        set.comps[set.count] = comps;
        ++set.count;
    }
    XK_TRY(launch_permute_set(set, n_watch, std::max<int64_t>(n_watch, 1), cap, (const int*)lp_watch.p, 1, st));
    if (!keep_on) lp_on = false;
    return 0;
}

int Ctx::grow_pairs(int64_t need) {
    if (need <= cap) return 0;
    XK_TRY(hipDeviceSynchronize());
    int64_t nc = std::max<int64_t>(need, std::max<int64_t>(1024, cap * 3 / 2));
    PairSoA old = P;
    std::vector<DBuf> nb(PB_COUNT);
    const size_t* sz = PB_SZ;
    for (int k = 0; k < PB_COUNT; ++k) XK_TRY(nb[k].ensure(sz[k] * (size_t)nc));
    // copy existing columns (component-major arrays are re-strided)
    auto cp = [&](int k, int comps, size_t esz) -> hipError_t {
        for (int c = 0; c < comps; ++c) {
            hipError_t e = hipMemcpy((char*)nb[k].p + c * nc * esz, (char*)pb[k].p + c * cap * esz,
                                     n_pairs * esz, hipMemcpyDeviceToDevice);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    if (n_pairs > 0) {
        for (int k = 0; k < PB_COUNT; ++k) {
            int comps = pb_comps(k);
            size_t esz = sz[k] / comps;
            XK_TRY(cp(k, comps, esz));
        }
    }
    for (int k = 0; k < PB_COUNT; ++k) {
        pb[k].release();
        pb[k] = nb[k];
    }
    (void)old;
    cap = nc;
    bind_pairs();
    return 0;
}

void Ctx::bind_pairs() {
    P.na0 = (int*)pb[PB_NA0].p; P.na1 = (int*)pb[PB_NA1].p;
    P.nb0 = (int*)pb[PB_NB0].p; P.nb1 = (int*)pb[PB_NB1].p;
    P.ra = (double*)pb[PB_RA].p; P.rb = (double*)pb[PB_RB].p;
    P.s = (double*)pb[PB_S].p; P.t = (double*)pb[PB_T].p; P.gap = (double*)pb[PB_GAP].p;
    P.n = (double*)pb[PB_N].p; P.t1 = (double*)pb[PB_T1].p;
    P.p_n = (double*)pb[PB_PN].p; P.k_pen = (double*)pb[PB_KPEN].p; P.k_t = (double*)pb[PB_KT].p;
    P.qtn = (double*)pb[PB_QTN].p; P.diss = (double*)pb[PB_DISS].p; P.z = (double*)pb[PB_Z].p;
    P.status = (int8_t*)pb[PB_STATUS].p; P.stick = (int8_t*)pb[PB_STICK].p;
    P.sinc = (int*)pb[PB_SINC].p;
    bind_coating(P, pb);
    P.cap = cap;
    ea = (int*)pb[PB_EA].p; eb = (int*)pb[PB_EB].p;
}

int Ctx::build_pattern() {
    XK_TRY(hipDeviceSynchronize());
    // a watch list belongs to the pattern it was selected on; its pairs' state goes back first
    if (lp_sync(nullptr, false) != 0) return -1;
    XK_TRY(hipDeviceSynchronize());
    solve_ready = false;
    tangent_staged = false;
    if (direct && !host_build) {
        if (build_pattern_gpu(nullptr) != 0) return -1;
        XK_TRY(d_nf.ensure((size_t)std::max(n_sinc, 1) * 12 * sizeof(double)));
        XK_TRY(d_sg.ensure((size_t)std::max(n_sinc, 1) * SG_REC * sizeof(double)));
        pattern_valid = true;
        return 0;
    }
    const int64_t np = n_pairs;
    std::vector<int8_t> st(np), cneed(coat_on ? np : 0);
    std::vector<double> pn(np), cc(coat_on ? np : 0), ccp(coat_on ? np : 0);
    std::vector<int> nd[4];
    if (np) {
        XK_TRY(hipMemcpy(st.data(), P.status, np, hipMemcpyDeviceToHost));
        XK_TRY(hipMemcpy(pn.data(), P.p_n, np * 8, hipMemcpyDeviceToHost));
        if (coat_on) {
            XK_TRY(hipMemcpy(cc.data(), P.cc, np * 8, hipMemcpyDeviceToHost));
            XK_TRY(hipMemcpy(ccp.data(), P.ccp, np * 8, hipMemcpyDeviceToHost));
            XK_TRY(hipMemcpy(cneed.data(), P.cneed, np, hipMemcpyDeviceToHost));
        }
        int* src[4] = {P.na0, P.na1, P.nb0, P.nb1};
        for (int k = 0; k < 4; ++k) {
            nd[k].resize(np);
            XK_TRY(hipMemcpy(nd[k].data(), src[k], np * 4, hipMemcpyDeviceToHost));
        }
    }
    // contributing pairs of the increment: status != INACTIVE or p_n > 0; with the
    // coating on also a compression (current, previous step or probed) > 0
    std::vector<int> sinc_of(np, -1);
    std::vector<int> sinc_list;
    for (int64_t i = 0; i < np; ++i)
        if (st[i] != ST_INACTIVE || pn[i] > 0.0 || (coat_on && (cc[i] > 0.0 || ccp[i] > 0.0 || cneed[i]))) {
            sinc_of[i] = (int)sinc_list.size();
            sinc_list.push_back((int)i);
        }
    const int ns = (int)sinc_list.size();
    const int ng = 2 * n_nodes;

    // block keys (row group, col group)
    std::vector<uint64_t> keys;
    keys.reserve(16 * ((size_t)n_elems + ns));
    auto key = [](uint64_t r, uint64_t c) { return (r << 32) | c; };
    for (int e = 0; e < n_elems; ++e) {
        int g[4] = {2 * conn_h[2 * e], 2 * conn_h[2 * e] + 1, 2 * conn_h[2 * e + 1], 2 * conn_h[2 * e + 1] + 1};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) keys.push_back(key(g[a], g[b]));
    }
    for (int q = 0; q < ns; ++q) {
        int i = sinc_list[q];
        int g[4] = {2 * nd[0][i], 2 * nd[1][i], 2 * nd[2][i], 2 * nd[3][i]};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) keys.push_back(key(g[a], g[b]));
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int nblk = (int)keys.size();
    gptr_h.assign(ng + 1, 0);
    gcol_h.resize(nblk);
    std::vector<int> brow(nblk);
    for (int b = 0; b < nblk; ++b) {
        int r = (int)(keys[b] >> 32);
        gcol_h[b] = (int)(keys[b] & 0xffffffffu);
        brow[b] = r;
        gptr_h[r + 1]++;
    }
    for (int r = 0; r < ng; ++r) gptr_h[r + 1] += gptr_h[r];
    auto find = [&](int r, int c) -> int {
        auto lo = gcol_h.begin() + gptr_h[r], hi = gcol_h.begin() + gptr_h[r + 1];
        auto it = std::lower_bound(lo, hi, c);
        return (int)(it - gcol_h.begin());
    };
    // block contribution lists
    std::vector<int> kb, cb, kp, cp;
    kb.reserve(16 * (size_t)n_elems);
    cb.reserve(16 * (size_t)n_elems);
    for (int e = 0; e < n_elems; ++e) {
        int g[4] = {2 * conn_h[2 * e], 2 * conn_h[2 * e] + 1, 2 * conn_h[2 * e + 1], 2 * conn_h[2 * e + 1] + 1};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                kb.push_back(find(g[a], g[b]));
                cb.push_back(e * 16 + a * 4 + b);
            }
    }
    std::vector<int> nk, nc;
    kp.reserve(16 * (size_t)ns);
    cp.reserve(16 * (size_t)ns);
    nk.reserve(4 * (size_t)ns);
    nc.reserve(4 * (size_t)ns);
    for (int q = 0; q < ns; ++q) {
        int i = sinc_list[q];
        int nn[4] = {nd[0][i], nd[1][i], nd[2][i], nd[3][i]};
        for (int a = 0; a < 4; ++a) {
            for (int b = 0; b < 4; ++b) {
                kp.push_back(find(2 * nn[a], 2 * nn[b]));
                cp.push_back(q * 16 + a * 4 + b);
            }
            nk.push_back(nn[a]);
            nc.push_back(q * 4 + a);
        }
    }
    std::vector<int> bptr, bsrc, pptr, psrc, nptr, nsrc;
    bucket(nblk, kb, cb, bptr, bsrc);
    bucket(nblk, kp, cp, pptr, psrc);
    bucket(n_nodes, nk, nc, nptr, nsrc);

    XK_TRY(upload(d_gptr, gptr_h));
    XK_TRY(upload(d_gcol, gcol_h));
    XK_TRY(upload(d_brow, brow));
    XK_TRY(upload(d_blk_bptr, bptr));
    XK_TRY(upload(d_blk_bsrc, bsrc));
    XK_TRY(upload(d_blk_pptr, pptr));
    XK_TRY(upload(d_blk_psrc, psrc));
    XK_TRY(upload(d_node_pptr, nptr));
    XK_TRY(upload(d_node_psrc, nsrc));
    {
        std::vector<int> qpos(std::max<size_t>(nsrc.size(), 1), 0);
        for (size_t q = 0; q < nsrc.size(); ++q) qpos[(size_t)nsrc[q]] = (int)q;
        XK_TRY(upload(d_node_qpos, qpos));
    }
    if (np) XK_TRY(hipMemcpy(P.sinc, sinc_of.data(), np * 4, hipMemcpyHostToDevice));
    if (direct) {
        // element -> block slots, and the blocks the gather must still visit
        XK_TRY(upload(d_eslot, kb));
        std::vector<int> desc;
        for (int b = 0; b < nblk; ++b) {
            if (pptr[b + 1] == pptr[b]) continue;
            const int r = brow[b], r0 = gptr_h[r], deg = gptr_h[r + 1] - r0;
            const int64_t base = 9 * (int64_t)r0 + 3 * (int64_t)(b - r0);
            if (base + 2 * 3 * (int64_t)deg + 3 > INT32_MAX) {
                set_error("tangent pattern exceeds 2^31 entries");
                return -1;
            }
            desc.push_back((int)base);
            desc.push_back(3 * deg);
            desc.push_back(pptr[b]);
            desc.push_back(pptr[b + 1] | (bptr[b + 1] > bptr[b] ? (int)0x80000000u : 0));
        }
        XK_TRY(upload(d_cdesc, desc));
        n_cdesc = (int)(desc.size() / 4);
    }
    XK_TRY(d_nf.ensure((size_t)std::max(ns, 1) * 12 * sizeof(double)));
    XK_TRY(d_sg.ensure((size_t)std::max(ns, 1) * SG_REC * sizeof(double)));
    n_sinc = ns;
    n_blocks = nblk;
    host_pattern_valid = true;
    pattern_valid = true;
    return 0;
}

int Ctx::host_pattern() {
    if (host_pattern_valid) return 0;
    gptr_h.resize(2 * (size_t)n_nodes + 1);
    gcol_h.resize(n_blocks);
    XK_TRY(hipMemcpy(gptr_h.data(), d_gptr.p, sizeof(int) * gptr_h.size(), hipMemcpyDeviceToHost));
    if (n_blocks) XK_TRY(hipMemcpy(gcol_h.data(), d_gcol.p, sizeof(int) * (size_t)n_blocks, hipMemcpyDeviceToHost));
    host_pattern_valid = true;
    return 0;
}

int Ctx::ensure_pattern() { return pattern_valid ? 0 : build_pattern(); }

BlockLists Ctx::block_lists() const {
    BlockLists B;
    B.gptr = (const int*)d_gptr.p; B.gcol = (const int*)d_gcol.p; B.brow = (const int*)d_brow.p;
    B.bptr = (const int*)d_blk_bptr.p; B.bsrc = (const int*)d_blk_bsrc.p;
    B.pptr = (const int*)d_blk_pptr.p; B.psrc = (const int*)d_blk_psrc.p;
    return B;
}

NodeLists Ctx::node_lists() const {
    NodeLists L;
    L.bptr = (const int*)d_node_bptr.p; L.bsrc = (const int*)d_node_bsrc.p;
    L.pptr = (const int*)d_node_pptr.p; L.psrc = (const int*)d_node_psrc.p;
    return L;
}

GatherArgs Ctx::gather_args(int node_mode) const {
    GatherArgs g{};
    g.block = gather_block;
    g.n_nodes = n_nodes;
    g.n_elems = n_elems;
    g.n_sinc = n_sinc;
    g.n_blocks = n_blocks;
    g.node_mode = node_mode;
    g.L = node_lists();
    g.B = block_lists();
    g.fe = (const double*)d_fe.p;
    g.nf = (const double*)d_nf.p;
    g.flags = (int*)d_flags.p;
    g.stop = nstop;
    if (direct) {
        g.fint = (const double*)d_fint.p;
        g.direct = 1;
        g.cdesc = (const int4*)d_cdesc.p;
        g.n_cdesc = n_cdesc;
    }
    return g;
}

hipError_t Ctx::launch_elements(const double* u, bool want_f, void* K, double mass_coef, const double* M,
                                hipStream_t st, int kprec) {
    if (direct)
        return launch_beam_direct(kprec >= 0 ? kprec : prec, n_elems, (const double*)d_X.p, u, (const int*)d_conn.p, bc,
                                  (const int8_t*)d_eflag.p, (const int*)d_eslot.p, (const int*)d_gptr.p,
                                  mass_coef, M, want_f ? (double*)d_fint.p : nullptr, K, (int*)d_flags.p, st,
                                  nstop);
    return launch_beam(prec, n_elems, (const double*)d_X.p, u, (const int*)d_conn.p, bc,
                       want_f ? (double*)d_fe.p : nullptr, K ? d_Ke.p : nullptr, (int*)d_flags.p, st, nstop);
}

}  // namespace xk

using namespace xk;

struct xk_ctx : public Ctx {};

namespace xk {

// core radii of the pair list from the per-element core radii
__global__ void coat_core_kernel(int64_t np, const int* __restrict__ ea, const int* __restrict__ eb,
                                 const double* __restrict__ crad, double* __restrict__ cra, double* __restrict__ crb) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) {
        cra[p] = crad[ea[p]];
        crb[p] = crad[eb[p]];
    }
}

// end of a converged step: compression_prev <- compression
__global__ void coat_commit_kernel(int64_t np, const double* __restrict__ cc, double* __restrict__ ccp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) ccp[p] = cc[p];
}

// #(pairs the probe marked that hold no slot of the pattern): one atomic per wave
__global__ void coat_unplaced_kernel(int64_t np, const int8_t* __restrict__ cneed, const int* __restrict__ sinc,
                                     int* __restrict__ count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool hit = p < np && cneed[p] != 0 && sinc[p] < 0;
    const unsigned long long m = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (int)__popcll(m));
}

// the same test on a running Newton call: a compressed pair without a slot
// moves the control record to XK_NC_REBUILD (every lane that finds one stores
// the same value; the attempt's later kernels then skip)
__global__ void coat_unplaced_ctl_kernel(double* __restrict__ ctl, int64_t np, const int8_t* __restrict__ cneed,
                                         const int* __restrict__ sinc) {
    if (!nc_go(ctl)) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np && cneed[p] != 0 && sinc[p] < 0) ctl[XK_NC_STATE] = XK_NC_REBUILD;
}

}  // namespace xk

extern "C" {

const char* xk_last_error(void) { return g_last_error.c_str(); }

int xk_ctx_create(xk_ctx** out, int n_nodes, int n_elems, const double* node_coords, const int32_t* connectivity,
                  const double* section8, double scf, const double* v_ref3, const double* elem_radius,
                  int precision) {
    if (!out || n_nodes <= 0 || n_elems < 0 || !node_coords || (n_elems && !connectivity) || !section8) {
        set_error("xk_ctx_create: invalid arguments");
        return -1;
    }
    const bool force_staged = (precision & XK_OPT_STAGED) != 0;
    const bool host_pattern = (precision & XK_OPT_HOST_PATTERN) != 0;
    precision &= ~(XK_OPT_STAGED | XK_OPT_HOST_PATTERN);
    if (precision != XK_PREC_FP64 && precision != XK_PREC_MIXED && precision != XK_PREC_FP32) {
        set_error("xk_ctx_create: unknown precision");
        return -1;
    }
    for (int i = 0; i < 2 * n_elems; ++i)
        if (connectivity[i] < 0 || connectivity[i] >= n_nodes) {
            set_error("xk_ctx_create: connectivity index out of range");
            return -1;
        }
    xk_ctx* c = new xk_ctx();
    c->n_nodes = n_nodes;
    c->n_elems = n_elems;
    c->prec = precision;
    c->host_build = host_pattern;
    c->X_h.assign(node_coords, node_coords + 3 * (size_t)n_nodes);
    c->conn_h.assign(connectivity, connectivity + 2 * (size_t)n_elems);
    c->rad_h.assign(n_elems, 0.0);
    if (elem_radius) c->rad_h.assign(elem_radius, elem_radius + n_elems);
    BeamConst& b = c->bc;
    b.E = section8[0]; b.G = section8[1]; b.A = section8[2]; b.Iy = section8[3];
    b.Iz = section8[4]; b.J = section8[5]; b.ky = section8[6]; b.kz = section8[7];
    b.scf = scf;
    b.has_vref = v_ref3 ? 1 : 0;
    for (int k = 0; k < 3; ++k) b.vref[k] = v_ref3 ? v_ref3[k] : 0.0;
    int rc = 0;
    auto fail = [&](const char* what) {
        set_error(std::string("xk_ctx_create: ") + what);
        rc = -1;
    };
    if (upload(c->d_X, c->X_h) != hipSuccess) fail("X");
    if (!rc && upload(c->d_conn, c->conn_h) != hipSuccess) fail("conn");
    if (!rc && upload(c->d_rad, c->rad_h) != hipSuccess) fail("radius");
    // core radii default to the contact radii (no coating layer)
    c->crad_h = c->rad_h;
    if (!rc && upload(c->d_crad, c->crad_h) != hipSuccess) fail("core radius");
    // static element -> node lists (element order)
    if (!rc) {
        std::vector<int> k, cd, ptr, src;
        for (int e = 0; e < n_elems; ++e)
            for (int a = 0; a < 2; ++a) {
                k.push_back(c->conn_h[2 * e + a]);
                cd.push_back(2 * e + a);
            }
        bucket(n_nodes, k, cd, ptr, src);
        if (upload(c->d_node_bptr, ptr) != hipSuccess || upload(c->d_node_bsrc, src) != hipSuccess)
            fail("node lists");
    }
    // element chains -> direct path
    if (!rc) {
        std::vector<int> deg(n_nodes, 0);
        bool chain = n_elems > 0;
        for (int e = 0; e < n_elems; ++e) {
            deg[c->conn_h[2 * e]]++;
            deg[c->conn_h[2 * e + 1]]++;
            // the direct element kernel stores each node's row run contiguously,
            // which needs element e to join nodes n and n + 1
            if (c->conn_h[2 * e + 1] != c->conn_h[2 * e] + 1) chain = false;
        }
        std::vector<int8_t> ef(std::max(n_elems, 1), 0);
        int shared = 0, deg2 = 0;
        for (int n = 0; n < n_nodes; ++n) {
            if (deg[n] > 2) chain = false;
            deg2 += deg[n] == 2;
        }
        for (int e = 0; e < n_elems; ++e) {
            const int n1 = c->conn_h[2 * e], n2 = c->conn_h[2 * e + 1];
            if (e > 0 && c->conn_h[2 * (e - 1) + 1] == n1 && deg[n1] == 2) ef[e] |= 1;
            if (e + 1 < n_elems && c->conn_h[2 * (e + 1)] == n2 && deg[n2] == 2) {
                ef[e] |= 2;
                ++shared;
            }
        }
        if (shared != deg2) chain = false;
        c->direct = chain && !force_staged;
        if (c->direct) {
            if (upload(c->d_eflag, ef) != hipSuccess) fail("element flags");
            // node flags of the device pattern build
            std::vector<int8_t> nf(n_nodes, 0);
            for (int e = 0; e < n_elems; ++e) {
                nf[c->conn_h[2 * e]] |= 3;
                nf[c->conn_h[2 * e + 1]] |= 1;
            }
            if (!rc && upload(c->d_nflag, nf) != hipSuccess) fail("node flags");
            // element chains: a new chain starts where no element joins n - 1 and n
            c->node_chain_h.assign(n_nodes, 0);
            for (int n = 0; n < n_nodes; ++n) {
                if (n == 0 || !(nf[n - 1] & 2)) c->cstart_h.push_back(n);
                c->node_chain_h[n] = (int)c->cstart_h.size() - 1;
            }
            c->cstart_h.push_back(n_nodes);
            c->n_chains = (int)c->cstart_h.size() - 1;
            if (!rc && (upload(c->d_cstart, c->cstart_h) != hipSuccess ||
                        upload(c->d_node_chain, c->node_chain_h) != hipSuccess))
                fail("element chains");
            if (!rc && (c->d_fint.ensure(sizeof(double) * 6 * (size_t)n_nodes) != hipSuccess ||
                        hipMemset(c->d_fint.p, 0, sizeof(double) * 6 * (size_t)n_nodes) != hipSuccess))
                fail("internal force buffer");
        }
    }
    if (!rc && c->d_fe.ensure((size_t)std::max(n_elems, 1) * 12 * 8) != hipSuccess) fail("fe stage");
    if (!rc && c->d_Ke.ensure((size_t)std::max(n_elems, 1) * 144 * (precision == 0 ? 8 : 4)) != hipSuccess)
        fail("Ke stage");
    if (!rc && c->d_flags.ensure(64) != hipSuccess) fail("flags");
    if (!rc && c->d_partial.ensure(sizeof(double) * (size_t)(gather_node_blocks(n_nodes, 128) + 1)) != hipSuccess)
        fail("partials");
    if (!rc && hipMemset(c->d_flags.p, 0, 64) != hipSuccess) fail("flags init");
    if (!rc && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) fail("side stream");
    if (!rc && (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess))
        fail("fork/join events");
    if (!rc) {
        c->pb.resize(PB_COUNT);
        if (c->grow_pairs(1024) != 0) rc = -1;
    }
    if (rc) {
        c->release_all();
        delete c;
        return -1;
    }
    *out = c;
    return 0;
}

int xk_ctx_destroy(xk_ctx* c) {
    if (!c) return 0;
    (void)hipDeviceSynchronize();
    c->release_all();
    delete c;
    return 0;
}

int xk_ctx_precision(const xk_ctx* c) { return c ? c->prec : -1; }

int xk_ctx_direct(const xk_ctx* c) { return c ? (c->direct ? 1 : 0) : -1; }

int xk_beam_assemble(xk_ctx* c, const double* u, double* f_int, void* K_vals, int64_t K_cap, void* stream) {
    if (!c || !u) {
        set_error("xk_beam_assemble: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (K_vals && c->ensure_pattern() != 0) return -1;
    if (K_vals && K_cap < 9 * (int64_t)c->n_blocks) {
        set_error("xk_beam_assemble: K_vals holds fewer entries than the current pattern");
        return -1;
    }
    XK_TRY(c->launch_elements(u, f_int != nullptr, K_vals, 0.0, nullptr, st));
    // node lanes copy / sum f_int; K lanes (direct path: contact blocks only) add nothing of the pairs
    GatherArgs g = c->gather_args(f_int ? 0 : -1);
    g.out = f_int;
    g.use_beam = 1;
    XK_TRY(launch_gather(c->prec, g, c->d_Ke.p, nullptr, K_vals, st));
    return 0;
}

int xk_pairs_count(const xk_ctx* c, int64_t* n) {
    if (!c || !n) return -1;
    *n = c->n_pairs;
    return 0;
}

int xk_pairs_set(xk_ctx* c, int64_t n, const int32_t* ea, const int32_t* eb, const double* s, const double* t,
                 const double* gap, const double* normal, const double* tangent1, const double* p_n,
                 const double* k_pen, const double* k_t, const double* q_trial_norm, const double* dissipation,
                 const double* z_t, const int8_t* status, const int8_t* stick) {
    if (!c || n < 0 || (n && (!ea || !eb))) {
        set_error("xk_pairs_set: invalid arguments");
        return -1;
    }
    for (int64_t i = 0; i < n; ++i)
        if (ea[i] < 0 || ea[i] >= c->n_elems || eb[i] < 0 || eb[i] >= c->n_elems) {
            set_error("xk_pairs_set: element index out of range");
            return -1;
        }
    XK_TRY(hipDeviceSynchronize());
    if (c->grow_pairs(std::max<int64_t>(n, 1)) != 0) return -1;
    c->n_pairs = n;
    c->lp_gprev_n = 0;
    c->lp_on = false;
    c->keys_valid = false;
    c->filtered_valid = false;
    if (n == 0) {
        c->pattern_valid = false;
        return 0;
    }
    const int64_t cap = c->cap;
    std::vector<int> na0(n), na1(n), nb0(n), nb1(n);
    std::vector<double> ra(n), rb(n), cra(n), crb(n);
    for (int64_t i = 0; i < n; ++i) {
        na0[i] = c->conn_h[2 * ea[i]]; na1[i] = c->conn_h[2 * ea[i] + 1];
        nb0[i] = c->conn_h[2 * eb[i]]; nb1[i] = c->conn_h[2 * eb[i] + 1];
        ra[i] = c->rad_h[ea[i]]; rb[i] = c->rad_h[eb[i]];
        cra[i] = c->crad_h[ea[i]]; crb[i] = c->crad_h[eb[i]];
    }
    auto put = [&](void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
    auto put_or_zero = [&](void* dst, const void* src, size_t bytes) {
        return src ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipMemset(dst, 0, bytes);
    };
    XK_TRY(put(c->ea, ea, n * 4));
    XK_TRY(put(c->eb, eb, n * 4));
    XK_TRY(put(c->P.na0, na0.data(), n * 4));
    XK_TRY(put(c->P.na1, na1.data(), n * 4));
    XK_TRY(put(c->P.nb0, nb0.data(), n * 4));
    XK_TRY(put(c->P.nb1, nb1.data(), n * 4));
    XK_TRY(put(c->P.ra, ra.data(), n * 8));
    XK_TRY(put(c->P.rb, rb.data(), n * 8));
    XK_TRY(put_or_zero(c->P.s, s, n * 8));
    XK_TRY(put_or_zero(c->P.t, t, n * 8));
    XK_TRY(put_or_zero(c->P.gap, gap, n * 8));
    for (int k = 0; k < 3; ++k) {
        std::vector<double> col(n, 0.0);
        if (normal) for (int64_t i = 0; i < n; ++i) col[i] = normal[3 * i + k];
        XK_TRY(put(c->P.n + k * cap, col.data(), n * 8));
        if (tangent1) for (int64_t i = 0; i < n; ++i) col[i] = tangent1[3 * i + k];
        else std::fill(col.begin(), col.end(), 0.0);
        XK_TRY(put(c->P.t1 + k * cap, col.data(), n * 8));
    }
    for (int k = 0; k < 2; ++k) {
        std::vector<double> col(n, 0.0);
        if (z_t) for (int64_t i = 0; i < n; ++i) col[i] = z_t[2 * i + k];
        XK_TRY(put(c->P.z + k * cap, col.data(), n * 8));
    }
    XK_TRY(put_or_zero(c->P.p_n, p_n, n * 8));
    XK_TRY(put_or_zero(c->P.k_pen, k_pen, n * 8));
    XK_TRY(put_or_zero(c->P.k_t, k_t, n * 8));
    XK_TRY(put_or_zero(c->P.qtn, q_trial_norm, n * 8));
    XK_TRY(put_or_zero(c->P.diss, dissipation, n * 8));
    XK_TRY(put_or_zero(c->P.status, status, n));
    if (stick) XK_TRY(put(c->P.stick, stick, n));
    else XK_TRY(hipMemset(c->P.stick, 1, n));
    XK_TRY(hipMemset(c->P.sinc, 0xff, n * 4));  // -1: not contributing until the pattern is built
    // coating: core radii of the elements, state of a new pair (xk_pairs_set_coating sets it)
    XK_TRY(put(c->P.cra, cra.data(), n * 8));
    XK_TRY(put(c->P.crb, crb.data(), n * 8));
    XK_TRY(hipMemset(c->P.cc, 0, n * 8));
    XK_TRY(hipMemset(c->P.ccp, 0, n * 8));
    XK_TRY(hipMemset(c->P.cz, 0, n * 8));
    XK_TRY(hipMemset(c->P.cz + cap, 0, n * 8));
    XK_TRY(hipMemset(c->P.cstick, 1, n));
    XK_TRY(hipMemset(c->P.cqtn, 0, n * 8));
    XK_TRY(hipMemset(c->P.cdiss, 0, n * 8));
    XK_TRY(hipMemset(c->P.cneed, 0, n));
    c->pattern_valid = false;
    return 0;
}

int xk_pairs_get(xk_ctx* c, int32_t* ea, int32_t* eb, double* s, double* t, double* gap, double* normal,
                 double* tangent1, double* p_n, double* k_pen, double* k_t, double* q_trial_norm,
                 double* dissipation, double* z_t, int8_t* status, int8_t* stick) {
    if (!c) return -1;
    XK_TRY(hipDeviceSynchronize());
    if (c->lp_sync(nullptr, false) != 0) return -1;
    XK_TRY(hipDeviceSynchronize());
    const int64_t n = c->n_pairs, cap = c->cap;
    if (n == 0) return 0;
    auto get = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        return dst ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
    };
    XK_TRY(get(ea, c->ea, n * 4));
    XK_TRY(get(eb, c->eb, n * 4));
    XK_TRY(get(s, c->P.s, n * 8));
    XK_TRY(get(t, c->P.t, n * 8));
    XK_TRY(get(gap, c->P.gap, n * 8));
    XK_TRY(get(p_n, c->P.p_n, n * 8));
    XK_TRY(get(k_pen, c->P.k_pen, n * 8));
    XK_TRY(get(k_t, c->P.k_t, n * 8));
    XK_TRY(get(q_trial_norm, c->P.qtn, n * 8));
    XK_TRY(get(dissipation, c->P.diss, n * 8));
    XK_TRY(get(status, c->P.status, n));
    XK_TRY(get(stick, c->P.stick, n));
    std::vector<double> col(n);
    for (int k = 0; k < 3; ++k) {
        if (normal) {
            XK_TRY(hipMemcpy(col.data(), c->P.n + k * cap, n * 8, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) normal[3 * i + k] = col[i];
        }
        if (tangent1) {
            XK_TRY(hipMemcpy(col.data(), c->P.t1 + k * cap, n * 8, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) tangent1[3 * i + k] = col[i];
        }
    }
    if (z_t)
        for (int k = 0; k < 2; ++k) {
            XK_TRY(hipMemcpy(col.data(), c->P.z + k * cap, n * 8, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) z_t[2 * i + k] = col[i];
        }
    return 0;
}

int xk_pairs_run(xk_ctx* c, int modes, const double* coords_ref, const double* u, const double* u_ref,
                 double k_pen, double smoothing_delta, double k_t_ratio, double mu_eff, double g_on, double g_off,
                 int allow_deactivation, void* stream) {
    if (!c) return -1;
    if (c->lp_sync((hipStream_t)stream, false) != 0) return -1;
    const int phys = PAIR_HUBER | PAIR_FRIC | PAIR_TAN_C | PAIR_TAN_F | PAIR_COAT;
    if (modes & PAIR_PROBE) {
        set_error("xk_pairs_run: the compression probe is xk_coating_probe");
        return -1;
    }
    if ((modes & PAIR_HYST) && (modes & phys)) {
        set_error("xk_pairs_run: the hysteresis update cannot be combined with force/tangent modes");
        return -1;
    }
    if ((modes & PAIR_HYST) && !(modes & PAIR_GEOM)) {
        set_error("xk_pairs_run: the hysteresis update needs the geometry mode");
        return -1;
    }
    if ((modes & PAIR_GEOM) && !coords_ref) {
        set_error("xk_pairs_run: geometry needs coordinates");
        return -1;
    }
    if ((modes & (PAIR_FRIC | PAIR_COAT_FRIC)) && !u) {
        set_error("xk_pairs_run: friction needs u");
        return -1;
    }
    if ((modes & PAIR_COAT) && !c->coat_on) {
        set_error("xk_pairs_run: the coating modes need the coating on (xk_ctx_set_coating)");
        return -1;
    }
    if ((modes & phys) && c->ensure_pattern() != 0) return -1;
    PairParams pp;
    pp.nqpos = (const int*)c->d_node_qpos.p;
    pp.k_pen = k_pen;
    pp.delta_h = smoothing_delta > 0.0 ? k_pen / smoothing_delta : 0.0;
    pp.k_t_ratio = k_t_ratio;
    pp.mu_eff = mu_eff;
    pp.mu_elem = c->has_mu_elem ? (const double*)c->d_mu_elem.p : nullptr;
    pp.ea = c->ea;
    pp.g_on = g_on;
    pp.g_off = g_off;
    pp.allow_deact = allow_deactivation;
    pp.n_sinc = c->n_sinc;
    c->coat_params(pp);
    XK_TRY(launch_pairs(c->prec, c->n_pairs, c->P, coords_ref, u, u_ref, modes, pp, (double*)c->d_nf.p,
                        c->d_sg.p, nullptr, (hipStream_t)stream, nullptr, c->pair_block));
    // a new compression can give an INACTIVE pair a coating contribution: the
    // pattern is rebuilt after every geometry update with the coating on
    if ((modes & PAIR_HYST) || (c->coat_on && (modes & PAIR_GEOM))) c->pattern_valid = false;
    c->last_modes = modes;
    return 0;
}

int xk_pairs_gather(xk_ctx* c, double* f, void* K_vals, int64_t K_cap, void* stream) {
    if (!c) return -1;
    if (c->ensure_pattern() != 0) return -1;
    if (K_vals && K_cap < 9 * (int64_t)c->n_blocks) {
        set_error("xk_pairs_gather: K_vals holds fewer entries than the current pattern");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (K_vals && c->direct) {
        // the direct gather visits contact blocks only: clear the rest first
        const size_t nnz = 9 * (size_t)c->n_blocks;
        XK_TRY(c->prec == 0 ? launch_fill_f64(nnz, (double*)K_vals, 0.0, st)
                            : launch_fill_f32(nnz, (float*)K_vals, 0.0f, st));
    }
    GatherArgs g = c->gather_args(f ? 1 : -1);
    g.out = f;
    g.use_pair = 1;
    XK_TRY(launch_gather(c->prec, g, nullptr, c->d_sg.p, K_vals, st));
    return 0;
}

int xk_pattern_build(xk_ctx* c) { return c ? c->build_pattern() : -1; }

int xk_newton_update(xk_ctx* c, double* u, double* du, const double* R, const double* f_contact,
                     const double* f_ext, const uint8_t* fixed_mask, const double* res_norm2, int line_search_steps,
                     double diverge_factor, double du_norm_cap, double* status, void* stream) {
    if (!c || !u || !du || !R || !f_contact || !f_ext || !fixed_mask || !res_norm2 || !status) {
        set_error("xk_newton_update: null argument");
        return -1;
    }
    return c->newton_update(u, du, R, f_contact, f_ext, fixed_mask, res_norm2, line_search_steps, diverge_factor,
                            du_norm_cap, status, (hipStream_t)stream);
}

int xk_newton_control_begin(xk_ctx* c, double* ctl, int hist_cap, double tol_force, double tol_disp,
                            double f_ext_ref_norm, int dynamic_ref, int divergence_window, int max_attempts,
                            double solve_tol, void* stream) {
    if (!c || !ctl) {
        set_error("xk_newton_control_begin: null argument");
        return -1;
    }
    return c->newton_control_begin(ctl, hist_cap, tol_force, tol_disp, f_ext_ref_norm, dynamic_ref,
                                   divergence_window, max_attempts, solve_tol, (hipStream_t)stream);
}

int xk_newton_control_attach(xk_ctx* c, double* ctl) {
    if (!c) {
        set_error("xk_newton_control_attach: null context");
        return -1;
    }
    c->nstop = ctl;
    return 0;
}

int xk_newton_control_force(xk_ctx* c, const double* res_norm2, const int32_t* n_active, void* stream) {
    if (!c || !res_norm2 || !n_active) {
        set_error("xk_newton_control_force: null argument");
        return -1;
    }
    return c->newton_control_force(res_norm2, n_active, (hipStream_t)stream);
}

int xk_newton_control_set(xk_ctx* c, int state, int reason, void* stream) {
    if (!c) {
        set_error("xk_newton_control_set: null context");
        return -1;
    }
    return c->newton_control_set(state, reason, (hipStream_t)stream);
}

int xk_solve_buffer(xk_ctx* c, int which, void* host, int64_t nbytes, int64_t* size) {
    return c ? c->solve_buffer(which, host, nbytes, size) : -1;
}

int xk_linear_solve(xk_ctx* c, const double* u, const uint8_t* fixed_mask, const xk_dyn* dyn, const double* R,
                    double* du, int max_refine, double tol, double* res_rel, int32_t* n_refine, void* stream) {
    if (!c || !u || !fixed_mask || !R || !du || max_refine < 0) {
        set_error("xk_linear_solve: invalid arguments");
        return -1;
    }
    double mass_coef = 0.0;
    const double* M = nullptr;
    if (dyn && dyn->dt > 1e-30) {
        const double c0 = 1.0 / (dyn->beta * (dyn->dt * dyn->dt));   // as the fused assembly forms it
        mass_coef = (1.0 - dyn->alpha_m) * c0;
        M = dyn->M_diag;
    }
    return c->linear_solve(u, fixed_mask, mass_coef, M, R, du, max_refine, tol, res_rel, n_refine,
                           (hipStream_t)stream);
}

int xk_solve_set_lu(xk_ctx* c, int kind, int32_t* pivoting, int64_t* fallbacks) {
    if (!c || kind < -1 || kind > 1) {
        set_error("xk_solve_set_lu: kind must be -1 (query), 0 or 1");
        return -1;
    }
    if (kind >= 0) c->sv_lu_kind = kind;
    if (pivoting) *pivoting = (c->sv_lu_kind == 1 || c->sv_gepp) ? 1 : 0;
    if (fallbacks) *fallbacks = c->sv_fallbacks;
    return 0;
}

int xk_pattern_info(xk_ctx* c, int64_t* nnz, int64_t* n_blocks, int64_t* n_contrib) {
    if (!c) return -1;
    if (c->ensure_pattern() != 0) return -1;
    if (nnz) *nnz = 9 * (int64_t)c->n_blocks;
    if (n_blocks) *n_blocks = c->n_blocks;
    if (n_contrib) *n_contrib = c->n_sinc;
    return 0;
}

int xk_pattern_export(xk_ctx* c, int64_t* indptr, int32_t* indices) {
    if (!c) return -1;
    if (c->ensure_pattern() != 0 || c->host_pattern() != 0) return -1;
    const int ng = 2 * c->n_nodes;
    int64_t pos = 0;
    for (int gr = 0; gr < ng; ++gr) {
        const int r0 = c->gptr_h[gr], r1 = c->gptr_h[gr + 1], deg = r1 - r0;
        for (int a = 0; a < 3; ++a) {
            const int64_t row = 3 * (int64_t)gr + a;
            if (indptr) indptr[row] = 9 * (int64_t)r0 + (int64_t)a * 3 * deg;
            if (indices)
                for (int b = r0; b < r1; ++b)
                    for (int k = 0; k < 3; ++k) indices[pos++] = 3 * c->gcol_h[b] + k;
        }
    }
    if (indptr) indptr[3 * (int64_t)ng] = 9 * (int64_t)c->n_blocks;
    return 0;
}

int xk_assemble_iteration(xk_ctx* c, const double* u, const double* u_ref, const double* f_ext,
                          const uint8_t* fixed_mask, const xk_dyn* dyn, double k_pen, double smoothing_delta,
                          double k_t_ratio, double mu_eff, double* R, void* K_vals, int64_t K_cap,
                          double* res_norm2, int32_t* n_active, void* stream) {
    if (!c || !u || !f_ext || !fixed_mask || !R) {
        set_error("xk_assemble_iteration: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (c->ensure_pattern() != 0) return -1;
    if (K_vals && K_cap < 9 * (int64_t)c->n_blocks) {
        set_error("xk_assemble_iteration: K_vals holds fewer entries than the current pattern");
        return -1;
    }
    PairParams pp;
    pp.nqpos = (const int*)c->d_node_qpos.p;
    pp.k_pen = k_pen;
    pp.delta_h = smoothing_delta > 0.0 ? k_pen / smoothing_delta : 0.0;
    pp.k_t_ratio = k_t_ratio;
    pp.mu_eff = mu_eff;
    pp.mu_elem = c->has_mu_elem ? (const double*)c->d_mu_elem.p : nullptr;
    pp.ea = c->ea;
    pp.g_on = 0.0;
    pp.g_off = 0.0;
    pp.allow_deact = 0;
    pp.n_sinc = c->n_sinc;
    pp.stop = c->nstop;
    c->coat_params(pp);
    if (c->coat_on && c->lp_on) {
        set_error("xk_assemble_iteration: the learned prescreen is not available with the coating on");
        return -1;
    }
    int modes = PAIR_GEOM | PAIR_HUBER | PAIR_FRIC | (K_vals ? (PAIR_TAN_C | PAIR_TAN_F) : 0);
    if (c->coat_on) modes |= PAIR_COAT_F | PAIR_COAT_FRIC | (K_vals ? (PAIR_COAT_K | PAIR_COAT_KF) : 0);
    c->tangent_staged = K_vals != nullptr;   // the pair kernel stages every contributing pair's tangent record
    // with the prescreen on, the pair kernel visits the watch list (every
    // contributing pair and the ones the network keeps), compacted
    const bool lp = c->lp_on;
    const int64_t n_work = lp ? c->n_watch : c->n_pairs;
    const PairSoA& PS = lp ? c->PW : c->P;
    if (lp) pp.ea = (const int*)c->wb[PB_EA].p;
    const int npb = pair_blocks(n_work, c->pair_block);
    XK_TRY(c->d_act.ensure(sizeof(int) * (size_t)std::max(npb, 1)));
    int* act = (int*)c->d_act.p;
    void* Ke = K_vals ? c->d_Ke.p : nullptr;
    GatherArgs g = c->gather_args(2);
    g.u = u;
    g.f_ext = f_ext;
    g.fixed = fixed_mask;
    g.out = R;
    g.use_beam = 1;
    g.use_pair = 1;
    if (dyn && dyn->dt > 1e-30) {
        if ((dyn->acc == nullptr) != (dyn->vel == nullptr)) {
            set_error("xk_assemble_iteration: dyn acc and vel are both set or both null");
            return -1;
        }
        DynArgs& dy = g.dy;
        dy.on = 1;
        dy.M = dyn->M_diag; dy.u_pred = dyn->u_pred; dy.v_pred = dyn->v_pred; dy.acc_old = dyn->acc_old;
        dy.acc = dyn->acc; dy.vel = dyn->vel;
        dy.c0 = 1.0 / (dyn->beta * (dyn->dt * dyn->dt));
        dy.dt = dyn->dt; dy.gamma = dyn->gamma; dy.alpha_m = dyn->alpha_m;
        g.mass_coef = (1.0 - dyn->alpha_m) * dy.c0;
        g.M = dyn->M_diag;
    }
    g.partial = (double*)c->d_partial.p;
    g.res_norm2 = res_norm2;
    g.act_part = act;
    g.n_act_part = n_work > 0 ? npb : 0;
    g.n_active = n_active;
    // the direct element kernel writes the beam blocks and the mass term into K
    const double emass = c->direct ? g.mass_coef : 0.0;
    hipEvent_t* ev = nullptr;
    if (c->prof_calls < c->prof_max) ev = &c->prof_ev[PROF_EV * (size_t)c->prof_calls++];
    if (ev) {
        // profiled: the three stages in sequence on the caller's stream
        XK_TRY(hipEventRecord(ev[0], st));
        XK_TRY(launch_pairs(c->prec, n_work, PS, (const double*)c->d_X.p, u, u_ref, modes, pp,
                            (double*)c->d_nf.p, c->d_sg.p, act, st, nullptr, c->pair_block));
        if (lp)
            XK_TRY(launch_lp_hist_push(c->n_nodes, u, (double*)c->lp_hist.p, Ctx::LP_HIST_CAP, (int*)c->lp_hist_cnt.p,
                                       st, c->nstop));
        XK_TRY(hipEventRecord(ev[1], st));
        XK_TRY(c->launch_elements(u, true, K_vals, emass, g.M, st));
        XK_TRY(hipEventRecord(ev[2], st));
    } else {
        // pair kernel on the side stream, concurrent with the element kernel
        // (this order measured 3% faster per step than the reverse)
        XK_TRY(hipEventRecord(c->ev_fork, st));
        XK_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        XK_TRY(launch_pairs(c->prec, n_work, PS, (const double*)c->d_X.p, u, u_ref, modes, pp,
                            (double*)c->d_nf.p, c->d_sg.p, act, c->side, nullptr, c->pair_block));
        // the geometry this assembly gives the pairs it skips, for the verification pass
        if (lp)
            XK_TRY(launch_lp_hist_push(c->n_nodes, u, (double*)c->lp_hist.p, Ctx::LP_HIST_CAP, (int*)c->lp_hist_cnt.p,
                                       c->side, c->nstop));
        XK_TRY(hipEventRecord(c->ev_join, c->side));
        XK_TRY(c->launch_elements(u, true, K_vals, emass, g.M, st));
        XK_TRY(hipStreamWaitEvent(st, c->ev_join, 0));
    }
    XK_TRY(launch_gather(c->prec, g, Ke, c->d_sg.p, K_vals, st, ev ? ev[3] : nullptr));
    if (ev) XK_TRY(hipEventRecord(ev[4], st));
    c->last_modes = modes;
    return 0;
}

int xk_profile_begin(xk_ctx* c, int max_calls) {
    if (!c || max_calls < 0) return -1;
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    c->prof_ev.assign(PROF_EV * (size_t)max_calls, nullptr);
    for (auto& e : c->prof_ev) XK_TRY(hipEventCreate(&e));
    c->prof_max = max_calls;
    c->prof_calls = 0;
    return 0;
}

int xk_profile_end(xk_ctx* c, double* stage_ms, int* n_calls) {
    if (!c) return -1;
    double acc[PROF_EV - 1] = {};
    const int n = c->prof_calls;
    if (n > 0) XK_TRY(hipEventSynchronize(c->prof_ev[PROF_EV * (size_t)(n - 1) + PROF_EV - 1]));
    for (int k = 0; k < n; ++k)
        for (int s = 0; s < PROF_EV - 1; ++s) {
            float ms = 0.f;
            XK_TRY(hipEventElapsedTime(&ms, c->prof_ev[PROF_EV * (size_t)k + s],
                                       c->prof_ev[PROF_EV * (size_t)k + s + 1]));
            acc[s] += ms;
        }
    if (stage_ms)
        for (int s = 0; s < PROF_EV - 1; ++s) stage_ms[s] = acc[s];
    if (n_calls) *n_calls = n;
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    c->prof_ev.clear();
    c->prof_max = 0;
    c->prof_calls = 0;
    return 0;
}

// ------------------------------------------------------------ learned prescreen
int xk_lp_set_params(xk_ctx* c, const float* params, int n_params, float threshold) {
    if (!c || !params || n_params != lp_n_params()) {
        set_error("xk_lp_set_params: expected " + std::to_string(lp_n_params()) + " parameters");
        return -1;
    }
    XK_TRY(c->lp_prm.ensure(sizeof(float) * (size_t)n_params));
    XK_TRY(hipMemcpy(c->lp_prm.p, params, sizeof(float) * (size_t)n_params, hipMemcpyHostToDevice));
    c->lp_thr = threshold;
    c->lp_ready = true;
    return 0;
}

int xk_lp_features(xk_ctx* c, const double* coords_ref, const double* u, const int32_t* elem_layer,
                   float layer_scale, float* features_out, void* stream) {
    if (!c || !coords_ref) {
        set_error("xk_lp_features: null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = c->n_pairs;
    if (n > c->lp_gprev_cap) {
        // grow the previous-gap column, keeping the entries of the pairs it already holds
        const int64_t nc = std::max<int64_t>(n, c->lp_gprev_cap * 3 / 2);
        DBuf nb;
        XK_TRY(nb.ensure(sizeof(double) * (size_t)nc));
        if (c->lp_gprev_n > 0)
            XK_TRY(hipMemcpyAsync(nb.p, c->lp_gprev.p, sizeof(double) * (size_t)c->lp_gprev_n,
                                  hipMemcpyDeviceToDevice, st));
        XK_TRY(hipStreamSynchronize(st));
        c->lp_gprev.release();
        c->lp_gprev = nb;
        c->lp_gprev_cap = nc;
    }
    XK_TRY(c->lp_feat.ensure(sizeof(uint16_t) * 8 * (size_t)std::max<int64_t>(n, 1)));
    XK_TRY(launch_lp_features(n, c->P, c->ea, c->eb, coords_ref, u, elem_layer, layer_scale,
                              (double*)c->lp_gprev.p, std::min(c->lp_gprev_n, n), (uint16_t*)c->lp_feat.p,
                              features_out, st));
    c->lp_gprev_n = n;
    return 0;
}

int xk_lp_select(xk_ctx* c, int64_t* n_watch, void* stream) {
    if (c && c->coat_on) {
        set_error("xk_lp_select: the learned prescreen is not available with the coating on");
        return -1;
    }
    if (!c) return -1;
    if (!c->lp_ready) {
        set_error("xk_lp_select: no parameters (xk_lp_set_params)");
        return -1;
    }
    if (c->lp_gprev_n != c->n_pairs) {
        set_error("xk_lp_select: features are not those of the current pair list (xk_lp_features)");
        return -1;
    }
    if (c->ensure_pattern() != 0) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = c->n_pairs;
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    XK_TRY(c->lp_keep.ensure(sizeof(int64_t) * nn));
    XK_TRY(c->lp_pos.ensure(sizeof(int64_t) * nn));
    XK_TRY(c->lp_watch.ensure(sizeof(int) * nn));
    XK_TRY(c->lp_prune.ensure(sizeof(int) * nn));
    XK_TRY(launch_lp_infer_bf16(n, (const uint16_t*)c->lp_feat.p, (const float*)c->lp_prm.p, c->lp_thr, c->P.sinc,
                                (int64_t*)c->lp_keep.p, nullptr, st));
    int64_t nw = 0;
    if (scan_exclusive_i64(c, 0, (const int64_t*)c->lp_keep.p, (int64_t*)c->lp_pos.p, n, &nw, st) != 0) return -1;
    XK_TRY(launch_lp_split(n, (const int64_t*)c->lp_keep.p, (const int64_t*)c->lp_pos.p, (int*)c->lp_watch.p,
                           (int*)c->lp_prune.p, st));
    c->n_watch = nw;
    c->n_prune = n - nw;
    XK_TRY(c->lp_hist.ensure(sizeof(double) * 3 * (size_t)c->n_nodes * Ctx::LP_HIST_CAP));
    XK_TRY(c->lp_hist_cnt.ensure(sizeof(int)));
    XK_TRY(launch_fill_i32(1, (int*)c->lp_hist_cnt.p, 0, st));
    if (c->lp_compact(st) != 0) return -1;
    XK_TRY(hipStreamSynchronize(st));
    c->lp_on = true;
    if (n_watch) *n_watch = nw;
    return 0;
}

int xk_lp_verify(xk_ctx* c, const double* coords_ref, const double* u, double g_on, int64_t* n_miss, void* stream) {
    if (!c || !coords_ref) return -1;
    (void)u;   // the geometries come from the assemblies' history
    hipStream_t st = (hipStream_t)stream;
    int64_t miss = 0;
    const bool was_on = c->lp_on;
    if (c->lp_sync(st, false) != 0) return -1;   // the watched pairs' state back in the pair list
    if (was_on && c->n_prune > 0) {
        // the pruned pairs get the narrowphase and frame transport of every
        // assembly of the call, in order (active set frozen, as in the call)
        int n_hist = 0;
        XK_TRY(hipMemcpyAsync(&n_hist, c->lp_hist_cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
        XK_TRY(hipStreamSynchronize(st));
        if (n_hist > Ctx::LP_HIST_CAP) miss = 1;   // history overflow: the call is re-run on the full path
        XK_TRY(launch_pair_replay(c->n_prune, (const int*)c->lp_prune.p, c->P, coords_ref, (const double*)c->lp_hist.p,
                                  3 * (int64_t)c->n_nodes, std::min(n_hist, (int)Ctx::LP_HIST_CAP), st));
        XK_TRY(launch_fill_i32(1, (int*)c->lp_hist_cnt.p, 0, st));
        XK_TRY(c->lp_count.ensure(sizeof(int)));
        XK_TRY(launch_fill_i32(1, (int*)c->lp_count.p, 0, st));
        XK_TRY(launch_lp_miss(c->n_prune, (const int*)c->lp_prune.p, c->P.gap, g_on, (int*)c->lp_count.p, st));
        int h = 0;
        XK_TRY(hipMemcpyAsync(&h, c->lp_count.p, sizeof(int), hipMemcpyDeviceToHost, st));
        XK_TRY(hipStreamSynchronize(st));
        miss += h;
    }
    c->lp_on = false;
    if (n_miss) *n_miss = miss;
    return 0;
}

int xk_lp_off(xk_ctx* c) {
    if (!c) return -1;
    XK_TRY(hipDeviceSynchronize());
    if (c->lp_sync(nullptr, false) != 0) return -1;
    XK_TRY(hipDeviceSynchronize());
    return 0;
}

int xk_pairs_snapshot(xk_ctx* c) {
    if (!c) return -1;
    XK_TRY(hipDeviceSynchronize());
    if (c->lp_sync(nullptr, true) != 0) return -1;   // current state, watch list kept
    XK_TRY(hipDeviceSynchronize());
    c->snap.resize(PB_COUNT);
    for (int k = 0; k < PB_COUNT; ++k) {
        XK_TRY(c->snap[k].ensure(std::max<size_t>(c->pb[k].bytes, 1)));
        if (c->pb[k].bytes)
            XK_TRY(hipMemcpy(c->snap[k].p, c->pb[k].p, c->pb[k].bytes, hipMemcpyDeviceToDevice));
    }
    c->snap_n = c->n_pairs;
    return 0;
}

int xk_pairs_restore(xk_ctx* c) {
    if (!c) return -1;
    if (c->snap_n != c->n_pairs || (int)c->snap.size() != PB_COUNT) {
        set_error("xk_pairs_restore: the pair list changed since the snapshot");
        return -1;
    }
    c->lp_on = false;   // the compact copy of the watched pairs is older than the restored state
    for (int k = 0; k < PB_COUNT; ++k) {
        if (c->snap[k].bytes < c->pb[k].bytes) {
            set_error("xk_pairs_restore: the pair list grew since the snapshot");
            return -1;
        }
        if (c->pb[k].bytes)
            XK_TRY(hipMemcpy(c->pb[k].p, c->snap[k].p, c->pb[k].bytes, hipMemcpyDeviceToDevice));
    }
    return 0;
}

int xk_mlp_n_params(void) { return lp_n_params(); }

int64_t xk_mlp_grad_work(int64_t n) { return (int64_t)lp_grad_blocks(n) * (lp_n_params() + 1); }

int xk_mlp_forward(const float* X, int64_t n, const float* params, float* logits, void* stream) {
    if (n < 0 || (n && (!X || !params || !logits))) {
        set_error("xk_mlp_forward: invalid arguments");
        return -1;
    }
    XK_TRY(launch_lp_infer_f32(n, X, params, 0.f, nullptr, logits, (hipStream_t)stream));
    return 0;
}

int xk_mlp_grad(const float* X, const float* y, int64_t n, const float* params, float pos_weight,
                float* grad_and_loss, float* work, int64_t work_len, void* stream) {
    if (n <= 0 || !X || !y || !params || !grad_and_loss || !work) {
        set_error("xk_mlp_grad: invalid arguments");
        return -1;
    }
    if (work_len < xk_mlp_grad_work(n)) {
        set_error("xk_mlp_grad: workspace too small (xk_mlp_grad_work)");
        return -1;
    }
    XK_TRY(launch_lp_grad(n, X, y, params, pos_weight, work, grad_and_loss, (hipStream_t)stream));
    return 0;
}

int xk_mlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr, float beta1, float beta2,
                float eps, void* stream) {
    if (!params || !grad || !m || !v || step < 1) {
        set_error("xk_mlp_adam: invalid arguments");
        return -1;
    }
    XK_TRY(launch_lp_adam(params, grad, m, v, step, lr, beta1, beta2, eps, (hipStream_t)stream));
    return 0;
}

int xk_ctx_set_friction(xk_ctx* c, const double* mu_elem) {
    if (!c) return -1;
    if (!mu_elem) {
        c->has_mu_elem = false;
        return 0;
    }
    for (int e = 0; e < c->n_elems; ++e)
        if (!(mu_elem[e] >= 0.0)) {
            set_error("xk_ctx_set_friction: coefficients must be >= 0");
            return -1;
        }
    XK_TRY(c->d_mu_elem.ensure(sizeof(double) * (size_t)std::max(c->n_elems, 1)));
    XK_TRY(hipMemcpy(c->d_mu_elem.p, mu_elem, sizeof(double) * (size_t)c->n_elems, hipMemcpyHostToDevice));
    c->has_mu_elem = true;
    return 0;
}

int xk_ctx_set_tuning(xk_ctx* c, const xk_tuning* t) {
    if (!c || !t) return -1;
    if (!(t->cell_factor > 0.0) || !pair_block_ok(t->pair_block) || !gather_block_ok(t->gather_block)) {
        set_error("xk_ctx_set_tuning: cell_factor > 0, pair_block in {64,128,256,512}, "
                  "gather_block in {128,256,512}");
        return -1;
    }
    c->cell_factor = t->cell_factor;
    c->pair_block = t->pair_block;
    c->gather_block = t->gather_block;
    return 0;
}

int xk_ctx_get_tuning(const xk_ctx* c, xk_tuning* t) {
    if (!c || !t) return -1;
    t->cell_factor = c->cell_factor;
    t->pair_block = c->pair_block;
    t->gather_block = c->gather_block;
    return 0;
}

int xk_ctx_set_coating(xk_ctx* c, double stiffness, double damping, double mu, double k_t_ratio,
                       const double* core_radii) {
    if (!c) return -1;
    if (!(damping >= 0.0) || !(mu >= 0.0) || !(k_t_ratio >= 0.0)) {
        set_error("xk_ctx_set_coating: damping, mu and k_t_ratio must be >= 0");
        return -1;
    }
    if (core_radii) {
        for (int e = 0; e < c->n_elems; ++e)
            if (!(core_radii[e] >= 0.0)) {
                set_error("xk_ctx_set_coating: core radii must be >= 0");
                return -1;
            }
        c->crad_h.assign(core_radii, core_radii + c->n_elems);
    } else {
        c->crad_h = c->rad_h;
    }
    XK_TRY(hipDeviceSynchronize());
    if (c->lp_sync(nullptr, false) != 0) return -1;
    XK_TRY(upload(c->d_crad, c->crad_h));
    if (c->n_pairs > 0) {
        hipLaunchKernelGGL(coat_core_kernel, dim3((unsigned)((c->n_pairs + 255) / 256)), dim3(256), 0, nullptr,
                           c->n_pairs, (const int*)c->ea, (const int*)c->eb, (const double*)c->d_crad.p, c->P.cra,
                           c->P.crb);
        XK_TRY(hipGetLastError());
        XK_TRY(hipDeviceSynchronize());
    }
    // _create_coating_strategy: Kelvin-Voigt when the stiffness is > 0, none otherwise
    c->coat_on = stiffness > 0.0;
    c->coat_k = stiffness;
    c->coat_c = damping;
    c->coat_mu = mu;
    c->coat_ktr = k_t_ratio;
    c->pattern_valid = false;
    return 0;
}

int xk_coating_set_dt(xk_ctx* c, double coat_dt) {
    if (!c) return -1;
    c->coat_dt = coat_dt;
    return 0;
}

int xk_coating_probe(xk_ctx* c, const double* u, int64_t* n_unplaced, void* stream) {
    if (!c || !n_unplaced) return -1;
    *n_unplaced = 0;
    if (!c->coat_on || c->n_pairs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (c->lp_sync(st, false) != 0) return -1;
    PairParams pp;
    pp.nqpos = (const int*)c->d_node_qpos.p;
    c->coat_params(pp);
    XK_TRY(launch_pairs(c->prec, c->n_pairs, c->P, (const double*)c->d_X.p, u, nullptr, PAIR_GEOM | PAIR_PROBE,
                        pp, (double*)c->d_nf.p, c->d_sg.p, nullptr, st, nullptr, c->pair_block));
    XK_TRY(c->d_coat_cnt.ensure(sizeof(int)));
    XK_TRY(hipMemsetAsync(c->d_coat_cnt.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(coat_unplaced_kernel, dim3((unsigned)((c->n_pairs + 255) / 256)), dim3(256), 0, st,
                       c->n_pairs, (const int8_t*)c->P.cneed, (const int*)c->P.sinc, (int*)c->d_coat_cnt.p);
    XK_TRY(hipGetLastError());
    int cnt = 0;
    XK_TRY(hipMemcpyAsync(&cnt, c->d_coat_cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
    XK_TRY(hipStreamSynchronize(st));
    *n_unplaced = cnt;
    if (cnt > 0) c->pattern_valid = false;   // the next pattern build gives the marked pairs a slot
    return 0;
}

int xk_coating_probe_control(xk_ctx* c, const double* u, void* stream) {
    if (!c) return -1;
    if (!c->nstop) {
        set_error("xk_coating_probe_control: no Newton control record attached");
        return -1;
    }
    if (!c->coat_on || c->n_pairs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (c->lp_sync(st, false) != 0) return -1;
    PairParams pp;
    pp.nqpos = (const int*)c->d_node_qpos.p;
    pp.stop = c->nstop;
    c->coat_params(pp);
    XK_TRY(launch_pairs(c->prec, c->n_pairs, c->P, (const double*)c->d_X.p, u, nullptr, PAIR_GEOM | PAIR_PROBE,
                        pp, (double*)c->d_nf.p, c->d_sg.p, nullptr, st, nullptr, c->pair_block));
    hipLaunchKernelGGL(coat_unplaced_ctl_kernel, dim3((unsigned)((c->n_pairs + 255) / 256)), dim3(256), 0, st,
                       c->nstop, c->n_pairs, (const int8_t*)c->P.cneed, (const int*)c->P.sinc);
    XK_TRY(hipGetLastError());
    return 0;
}

int xk_coating_commit(xk_ctx* c, void* stream) {
    if (!c) return -1;
    if (c->n_pairs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (c->lp_sync(st, false) != 0) return -1;
    hipLaunchKernelGGL(coat_commit_kernel, dim3((unsigned)((c->n_pairs + 255) / 256)), dim3(256), 0, st,
                       c->n_pairs, (const double*)c->P.cc, c->P.ccp);
    XK_TRY(hipGetLastError());
    return 0;
}

int xk_pairs_get_coating(xk_ctx* c, double* core_a, double* core_b, double* compression,
                         double* compression_prev, double* z_t, int8_t* stick, double* q_trial_norm,
                         double* dissipation) {
    if (!c) return -1;
    XK_TRY(hipDeviceSynchronize());
    if (c->lp_sync(nullptr, false) != 0) return -1;
    XK_TRY(hipDeviceSynchronize());
    const int64_t n = c->n_pairs, cap = c->cap;
    if (n == 0) return 0;
    auto get = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        return dst ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
    };
    XK_TRY(get(core_a, c->P.cra, n * 8));
    XK_TRY(get(core_b, c->P.crb, n * 8));
    XK_TRY(get(compression, c->P.cc, n * 8));
    XK_TRY(get(compression_prev, c->P.ccp, n * 8));
    XK_TRY(get(stick, c->P.cstick, n));
    XK_TRY(get(q_trial_norm, c->P.cqtn, n * 8));
    XK_TRY(get(dissipation, c->P.cdiss, n * 8));
    if (z_t) {
        std::vector<double> col(n);
        for (int k = 0; k < 2; ++k) {
            XK_TRY(hipMemcpy(col.data(), c->P.cz + k * cap, n * 8, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) z_t[2 * i + k] = col[i];
        }
    }
    return 0;
}

int xk_pairs_set_coating(xk_ctx* c, const double* compression, const double* compression_prev, const double* z_t,
                         const int8_t* stick, const double* q_trial_norm, const double* dissipation) {
    if (!c) return -1;
    XK_TRY(hipDeviceSynchronize());
    if (c->lp_sync(nullptr, false) != 0) return -1;
    const int64_t n = c->n_pairs, cap = c->cap;
    if (n == 0) return 0;
    auto put = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        return src ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    XK_TRY(put(c->P.cc, compression, n * 8));
    XK_TRY(put(c->P.ccp, compression_prev, n * 8));
    XK_TRY(put(c->P.cstick, stick, n));
    XK_TRY(put(c->P.cqtn, q_trial_norm, n * 8));
    XK_TRY(put(c->P.cdiss, dissipation, n * 8));
    if (z_t) {
        std::vector<double> col(n);
        for (int k = 0; k < 2; ++k) {
            for (int64_t i = 0; i < n; ++i) col[i] = z_t[2 * i + k];
            XK_TRY(hipMemcpy(c->P.cz + k * cap, col.data(), n * 8, hipMemcpyHostToDevice));
        }
    }
    c->pattern_valid = false;
    return 0;
}

int xk_status(xk_ctx* c, int* flags) {
    if (!c || !flags) return -1;
    XK_TRY(hipDeviceSynchronize());
    XK_TRY(hipMemcpy(flags, c->d_flags.p, sizeof(int), hipMemcpyDeviceToHost));
    XK_TRY(hipMemset(c->d_flags.p, 0, sizeof(int)));
    return 0;
}

}  // extern "C"
