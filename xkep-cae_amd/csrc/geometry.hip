// Constraint Jacobian of the point-to-point contact geometry, on the device:
// G = d g_n / d u, one row per pair whose status is not INACTIVE, in pair
// order (_build_constraint_jacobian_ptp, xkep_cae/contact/geometry/strategy.py:83-132).
//
// Row r (pair i): entries c_k n_d at the translational DOFs node_k * ndpn + d
// of the pair's four nodes [a0, a1, b0, b1] with c = [(1 - s), s, -(1 - t), -t];
// an entry with |c_k n_d| <= 1e-30 is not stored, and entries that fall on one
// DOF (a node shared by both elements) are summed -- the CSR that SciPy's
// COO -> CSR conversion (duplicates summed, columns sorted) makes of the
// reference's triplets.  At most two triplets meet on a DOF (a0 != a1 and
// b0 != b1), so the sum does not depend on its order.
//
// Two passes: per-pair row lengths and the active flag, exclusive scans for
// the row offsets and the row numbers, then the rows.

#include <algorithm>
#include <string>

#include "xk_common.h"
#include "xk_structs.h"
#include "xk_internal.h"
#include "../../include/xkamd.h"

namespace xk {

struct JacRow {
    int n;           // stored entries
    int col[12];
    double val[12];
};

__device__ __forceinline__ void jac_row(const PairSoA& P, int64_t i, int ndpn, JacRow& row) {
    const int64_t cap = P.cap;
    const double s = P.s[i], t = P.t[i];
    const double nv[3] = {P.n[i], P.n[cap + i], P.n[2 * cap + i]};
    const double c[4] = {1.0 - s, s, -(1.0 - t), -t};
    int node[4] = {P.na0[i], P.na1[i], P.nb0[i], P.nb1[i]};
    // node order with equal nodes merged; the coefficient of each node kept per
    // original slot so the drop rule applies to every triplet before the sum
    int ord[4] = {0, 1, 2, 3};
    for (int a = 1; a < 4; ++a) {   // insertion sort of 4 slots by node id
        const int v = ord[a];
        int b = a - 1;
        while (b >= 0 && node[ord[b]] > node[v]) {
            ord[b + 1] = ord[b];
            --b;
        }
        ord[b + 1] = v;
    }
    row.n = 0;
    for (int q = 0; q < 4;) {
        const int nd = node[ord[q]];
        int q2 = q + 1;
        while (q2 < 4 && node[ord[q2]] == nd) ++q2;
        for (int d = 0; d < 3; ++d) {
            bool any = false;
            double sum = 0.0;
            for (int r = q; r < q2; ++r) {
                const double v = c[ord[r]] * nv[d];
                if (fabs(v) > 1e-30) {
                    sum = any ? sum + v : v;
                    any = true;
                }
            }
            if (any) {
                row.col[row.n] = nd * ndpn + d;
                row.val[row.n] = sum;
                ++row.n;
            }
        }
        q = q2;
    }
}

__global__ void jac_count_kernel(int64_t n, PairSoA P, int ndpn, int64_t* __restrict__ act,
                                 int64_t* __restrict__ len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (P.status[i] == ST_INACTIVE) {
        act[i] = 0;
        len[i] = 0;
        return;
    }
    JacRow row;
    jac_row(P, i, ndpn, row);
    act[i] = 1;
    len[i] = row.n;
}

__global__ void jac_write_kernel(int64_t n, PairSoA P, int ndpn, const int64_t* __restrict__ act,
                                 const int64_t* __restrict__ rowpos, const int64_t* __restrict__ off,
                                 int32_t* __restrict__ active_idx, int64_t* __restrict__ indptr,
                                 int32_t* __restrict__ cols, double* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !act[i]) return;
    JacRow row;
    jac_row(P, i, ndpn, row);
    const int64_t r = rowpos[i], o = off[i];
    active_idx[r] = (int32_t)i;
    indptr[r] = o;
    for (int k = 0; k < row.n; ++k) {
        cols[o + k] = row.col[k];
        vals[o + k] = row.val[k];
    }
}

}  // namespace xk

using namespace xk;

struct xk_ctx : public Ctx {};

#define GJ_TRY(expr)                                                                 \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return -1;                                                               \
        }                                                                            \
    } while (0)

extern "C" int xk_pairs_constraint_jacobian(xk_ctx* c, int ndof_per_node, int64_t ndof_total, int64_t* n_rows,
                                            int64_t* nnz, int32_t* active_idx, int64_t* indptr, int32_t* indices,
                                            double* data, int64_t cap_rows, int64_t cap_nnz, void* stream) {
    if (!c || !n_rows || !nnz || ndof_per_node < 3) {
        set_error("xk_pairs_constraint_jacobian: invalid arguments");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (c->lp_sync(st, true) != 0) return -1;
    const int64_t np = c->n_pairs;
    *n_rows = 0;
    *nnz = 0;
    if (np == 0) return 0;
    if ((int64_t)c->n_nodes * ndof_per_node > ndof_total) {
        set_error("xk_pairs_constraint_jacobian: ndof_total is smaller than n_nodes x ndof_per_node");
        return -1;
    }
    GJ_TRY(c->gj.ensure(sizeof(int64_t) * 4 * (size_t)np));
    int64_t* act = (int64_t*)c->gj.p;
    int64_t* len = act + np;
    int64_t* rowpos = len + np;
    int64_t* off = rowpos + np;
    const unsigned nb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(jac_count_kernel, dim3(nb), dim3(256), 0, st, np, c->P, ndof_per_node, act, len);
    GJ_TRY(hipGetLastError());
    int64_t rows = 0, total = 0;
    if (scan_exclusive_i64(c, 0, act, rowpos, np, &rows, st) != 0) return -1;
    if (scan_exclusive_i64(c, 0, len, off, np, &total, st) != 0) return -1;
    *n_rows = rows;
    *nnz = total;
    if (!active_idx && !indptr && !indices && !data) return 0;   // size query
    if (!active_idx || !indptr || !indices || !data) {
        set_error("xk_pairs_constraint_jacobian: output arrays must all be given");
        return -1;
    }
    if (cap_rows < rows || cap_nnz < total) {
        set_error("xk_pairs_constraint_jacobian: output capacity too small");
        return -1;
    }
    hipLaunchKernelGGL(jac_write_kernel, dim3(nb), dim3(256), 0, st, np, c->P, ndof_per_node, act, rowpos, off,
                       active_idx, indptr, indices, data);
    GJ_TRY(hipGetLastError());
    GJ_TRY(hipMemcpyAsync(indptr + rows, &total, sizeof(int64_t), hipMemcpyHostToDevice, st));
    GJ_TRY(hipStreamSynchronize(st));
    return 0;
}
