// Shared device/host definitions of the gfx950 assembly library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define XK_HD __host__ __device__ __forceinline__

namespace xk {

template <class T>
struct V3 {
    T x, y, z;
};

template <class T>
XK_HD V3<T> v3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T>
XK_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
XK_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
XK_HD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
XK_HD V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
XK_HD T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T>
XK_HD V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
XK_HD T norm(V3<T> a) { return sqrt(dot(a, a)); }

// Row-major 3x3.
template <class T>
struct M3 {
    T m[9];
    XK_HD T& operator()(int i, int j) { return m[3 * i + j]; }
    XK_HD T operator()(int i, int j) const { return m[3 * i + j]; }
};

template <class T>
XK_HD M3<T> eye3() {
    M3<T> r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.m[k] = (k % 4 == 0) ? T(1) : T(0);
    return r;
}
template <class T>
XK_HD M3<T> skew(V3<T> v) {
    M3<T> S;
    S.m[0] = T(0); S.m[1] = -v.z; S.m[2] = v.y;
    S.m[3] = v.z;  S.m[4] = T(0); S.m[5] = -v.x;
    S.m[6] = -v.y; S.m[7] = v.x;  S.m[8] = T(0);
    return S;
}
template <class T>
XK_HD M3<T> mul(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = (a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j]) + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}
// a^T b
template <class T>
XK_HD M3<T> mulTN(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = (a.m[i] * b.m[j] + a.m[3 + i] * b.m[3 + j]) + a.m[6 + i] * b.m[6 + j];
    return r;
}
// a b^T
template <class T>
XK_HD M3<T> mulNT(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = (a.m[3 * i] * b.m[3 * j] + a.m[3 * i + 1] * b.m[3 * j + 1]) + a.m[3 * i + 2] * b.m[3 * j + 2];
    return r;
}
template <class T>
XK_HD V3<T> mulv(const M3<T>& a, V3<T> v) {
    return {(a.m[0] * v.x + a.m[1] * v.y) + a.m[2] * v.z,
            (a.m[3] * v.x + a.m[4] * v.y) + a.m[5] * v.z,
            (a.m[6] * v.x + a.m[7] * v.y) + a.m[8] * v.z};
}
// a^T v
template <class T>
XK_HD V3<T> mulTv(const M3<T>& a, V3<T> v) {
    return {(a.m[0] * v.x + a.m[3] * v.y) + a.m[6] * v.z,
            (a.m[1] * v.x + a.m[4] * v.y) + a.m[7] * v.z,
            (a.m[2] * v.x + a.m[5] * v.y) + a.m[8] * v.z};
}
template <class T, class S>
XK_HD M3<T> cvt(const M3<S>& a) {
    M3<T> r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.m[k] = T(a.m[k]);
    return r;
}
template <class T, class S>
XK_HD V3<T> cvt(V3<S> a) { return {T(a.x), T(a.y), T(a.z)}; }

// Device status word bits (mirror of XK_FLAG_* in include/xkamd.h).
constexpr int FLAG_NONFINITE = 1;
constexpr int FLAG_DEGENERATE = 2;
constexpr int FLAG_COAT_UNPLACED = 8;

// Pair-kernel modes (mirror of XK_PAIR_*).
constexpr int PAIR_GEOM = 1;
constexpr int PAIR_HYST = 2;
constexpr int PAIR_HUBER = 4;
constexpr int PAIR_FRIC = 8;
constexpr int PAIR_TAN_C = 16;
constexpr int PAIR_TAN_F = 32;
constexpr int PAIR_PROBE = 64;
constexpr int PAIR_COAT_F = 128;      // coating normal force (Kelvin-Voigt)
constexpr int PAIR_COAT_FRIC = 256;   // coating Coulomb return mapping and force
constexpr int PAIR_COAT_K = 512;      // coating normal stiffness
constexpr int PAIR_COAT_KF = 1024;    // coating friction stiffness
constexpr int PAIR_COAT = PAIR_COAT_F | PAIR_COAT_FRIC | PAIR_COAT_K | PAIR_COAT_KF;

// ContactStatus codes (xkep_cae/contact/_types.py:8-11).
constexpr int8_t ST_INACTIVE = 0;
constexpr int8_t ST_ACTIVE = 1;
constexpr int8_t ST_SLIDING = 2;

// Beam section constants passed by value to the element kernel.
struct BeamConst {
    double E, G, A, Iy, Iz, J, ky, kz;
    double scf;      // <= 0: no slenderness correction
    int has_vref;
    double vref[3];
};

// Per contributing pair, the pair kernel stages its force at its four nodes,
// c_k F (3 FP64 values each), in node order -- at the position of (pair slot,
// k) in the node contribution lists, so the gather sums contiguous runs --
// and one tangent record [G00 .. G22, c0, c1, c2, c3, pad x3] (FP64; 16
// entries, two aligned 64-byte segments); element records are 12 FP64 forces
// and 144 tangent values.
constexpr int SG_REC = 16;
constexpr int SG_C = 9;   // offset of c[4] in the tangent record

// Workgroups of one launch are dispatched round-robin over the 8 XCDs of the
// device (workgroup b on XCD b % 8), each XCD with its own L2.  For a grid
// padded to a multiple of 8, logical block (b % 8) * (grid / 8) + b / 8 gives
// every XCD one contiguous range of logical blocks, so work items with nearby
// indices (pairs of nearby elements, rows of nearby nodes) share an L2.
constexpr int N_XCD = 8;
__host__ __device__ inline int xcd_padded(int n_blocks) { return (n_blocks + N_XCD - 1) / N_XCD * N_XCD; }
__device__ __forceinline__ int xcd_logical_block() {
    const int per = (int)gridDim.x / N_XCD;
    return ((int)blockIdx.x % N_XCD) * per + (int)blockIdx.x / N_XCD;
}

// Device Newton control: the state word of the control record of the call
// in progress (null: no call attached).  The kernels of an attempt run only
// while the call runs (state 0); an attempt enqueued after the call ended,
// or while the host re-solves an attempt, is a no-op.  The value is uniform
// over the grid.
__device__ __forceinline__ bool nc_go(const double* stop) { return !stop || *stop == 0.0; }

}  // namespace xk
