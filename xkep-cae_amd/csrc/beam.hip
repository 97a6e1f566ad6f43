// Corotational Timoshenko beam element: internal force and symmetrised
// analytical tangent, one element per lane.
//
// Follows the reference's batched element math
// (xkep_cae/elements/_beam_assembly.py:211-325 with the helpers of
// xkep_cae/elements/_beam_cr.py:672-998).  Kinematics (local axes, minimal
// rotation, exponential/logarithmic rotation maps), the corotated
// displacements and the element force are FP64 in every precision mode; the
// tangent (B operator, Ke B, material and geometric parts, symmetrisation) is
// evaluated in TT = double (plain) or float (mixed).
//
// The tangent is formed without the dense 12x12 products: the nonzero rows
// of B are {3,4,5,6,9,10,11} and every row of Ke B and of K_geo has the
// column-block form [x, y, -x, z] (x: node-1 translations, y: node-1
// rotations, z: node-2 rotations), so each 3-row block b of the unsymmetrised
// tangent is carried as three 3x3 matrices X_b, Y_b, Z_b.
//
// Two output paths share the element math (beam_core):
//  * direct (meshes made of element chains with consecutive node numbers,
//    every strand): a workgroup takes 64 consecutive elements and owns the
//    first 63, i.e. the node 2 of each.  The internal force of an owned node
//    and its tangent rows (the element's node-2 blocks plus the next
//    element's node-1 blocks, exchanged through LDS in element order, plus
//    the lumped-mass term) go straight into f_int and K, the rows as
//    contiguous runs stored by consecutive lanes; the 64th lane only supplies
//    its node-1 part (it is owned by the next workgroup).  Nothing is staged.
//  * staged (any connectivity): per-element records
//    fe_stage[12 e + i] (global-frame force) and
//    Ke_stage[144 e + 9 lb + 3 a + b] (symmetrised 3x3 group blocks,
//    lb = g_row * 4 + g_col over {node1 trans, node1 rot, node2 trans,
//    node2 rot}), written through LDS as contiguous runs; the gather kernel
//    sums them per node / per block.

#include "xk_common.h"

#include <type_traits>

namespace xk {

// Half-angle of a rotation: angle a, sin(a/2), cos(a/2) (the tangent
// operators of the FP32 tangent reuse them instead of trigonometry of their own).
struct HalfAngle {
    double a, s, c;
};

// Unit quaternion (w, v) of a rotation vector, Taylor below |th| < 1e-12
// (_beam_cr.py:685-726), returned as the rotation matrix.
__device__ __forceinline__ M3<double> rotvec_to_rotmat(V3<double> th, HalfAngle* ha = nullptr) {
    double ang = norm(th);
    double half = ang / 2.0;
    double sinc, ch;
    if (ang < 1e-12) {
        double h2 = half * half;
        sinc = 0.5 - h2 / 48.0;
        ch = 1.0 - h2 / 2.0;
    } else {
        double sh;
        sincos(half, &sh, &ch);   // one argument reduction for both
        sinc = sh / ang;
    }
    if (ha) *ha = HalfAngle{ang, sinc * ang, ch};
    double w = ch, x = sinc * th.x, y = sinc * th.y, z = sinc * th.z;
    double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
    double wx = w * x, wy = w * y, wz = w * z;
    M3<double> R;
    R.m[0] = 1.0 - 2.0 * (yy + zz); R.m[1] = 2.0 * (xy - wz);       R.m[2] = 2.0 * (xz + wy);
    R.m[3] = 2.0 * (xy + wz);       R.m[4] = 1.0 - 2.0 * (xx + zz); R.m[5] = 2.0 * (yz - wx);
    R.m[6] = 2.0 * (xz - wy);       R.m[7] = 2.0 * (yz + wx);       R.m[8] = 1.0 - 2.0 * (xx + yy);
    return R;
}

// Logarithmic map by Shepperd's method: branch = argmax(tr, R00, R11, R22)
// (first on ties), w >= 0, normalise (floor 1e-15), 2 atan2(|v|, w)/|v| v,
// 2 v below |v| < 1e-12 (_beam_cr.py:729-787).
// FAST: the divisions by s and by |q| are one reciprocal and products (FP32
// tangent form; the FP64 form divides as the reference does).
template <bool FAST = false>
__device__ __forceinline__ V3<double> rotmat_to_rotvec(const M3<double>& R, HalfAngle* ha = nullptr) {
    double tr = R.m[0] + R.m[4] + R.m[8];
    int best = 0;
    double bv = tr;
    if (R.m[0] > bv) { best = 1; bv = R.m[0]; }
    if (R.m[4] > bv) { best = 2; bv = R.m[4]; }
    if (R.m[8] > bv) { best = 3; }
    // branch b: s = 2 sqrt(arg), the component b is s / 4, the others num_k / s
    double arg, n0, n1, n2;
    if (best == 0) {
        arg = tr + 1.0;
        n0 = R.m[7] - R.m[5]; n1 = R.m[2] - R.m[6]; n2 = R.m[3] - R.m[1];
    } else if (best == 1) {
        arg = ((1.0 + R.m[0]) - R.m[4]) - R.m[8];
        n0 = R.m[7] - R.m[5]; n1 = R.m[1] + R.m[3]; n2 = R.m[2] + R.m[6];
    } else if (best == 2) {
        arg = ((1.0 + R.m[4]) - R.m[0]) - R.m[8];
        n0 = R.m[2] - R.m[6]; n1 = R.m[1] + R.m[3]; n2 = R.m[5] + R.m[7];
    } else {
        arg = ((1.0 + R.m[8]) - R.m[0]) - R.m[4];
        n0 = R.m[3] - R.m[1]; n1 = R.m[2] + R.m[6]; n2 = R.m[5] + R.m[7];
    }
    const double s = 2.0 * sqrt(arg);
    double d0, d1, d2;
    if (FAST) {
        const double is = 1.0 / s;
        d0 = n0 * is; d1 = n1 * is; d2 = n2 * is;
    } else {
        d0 = n0 / s; d1 = n1 / s; d2 = n2 / s;
    }
    const double qb = 0.25 * s;
    double q0, q1, q2, q3;
    if (best == 0) { q0 = qb; q1 = d0; q2 = d1; q3 = d2; }
    else if (best == 1) { q0 = d0; q1 = qb; q2 = d1; q3 = d2; }
    else if (best == 2) { q0 = d0; q1 = d1; q2 = qb; q3 = d2; }
    else { q0 = d0; q1 = d1; q2 = d2; q3 = qb; }
    if (q0 < 0.0) { q0 = -q0; q1 = -q1; q2 = -q2; q3 = -q3; }
    double qn = sqrt(((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3);
    qn = qn > 1e-15 ? qn : 1e-15;
    if (FAST) {
        const double iq = 1.0 / qn;
        q0 *= iq; q1 *= iq; q2 *= iq; q3 *= iq;
    } else {
        q0 /= qn; q1 /= qn; q2 /= qn; q3 /= qn;
    }
    V3<double> v{q1, q2, q3};
    double sh = norm(v);
    double coef = (sh < 1e-12) ? 2.0 : 2.0 * atan2(sh, q0) / sh;
    if (ha) *ha = HalfAngle{coef * sh, sh, q0};
    return coef * v;
}

// T_s(th) and its inverse (_beam_cr.py:856-908) in FP32 from the half-angle
// sine and cosine: c1 = (1 - cos a)/a^2 = 2 (sin(a/2)/a)^2; c2 = (1 - sin a/a)/a^2
// and, for the inverse, 1/a^2 - (1 + cos a)/(2 a sin a) = 1/a^2 - cos(a/2)/(2 a sin(a/2)),
// by their Taylor series below a = 1/2 (truncation < 1e-12) where the closed
// forms cancel in FP32; the reference's own constants below a = 1e-10.
// The operator is I + c1 S(th) + c2 S(th)^2; its coefficients are formed where
// the rotation maps give the half angle (so the FP64 half-angle values are not
// kept until the tangent), the matrix at the tangent.
struct TanOpF {
    V3<float> th;
    float c1, c2;
};

__device__ __forceinline__ TanOpF tangent_coef_f32(V3<double> thd, const HalfAngle& h, bool inverse) {
    const float a = (float)h.a, s = (float)h.s, c = (float)h.c;
    const float a2 = a * a;
    float c1, c2;
    if (!inverse) {
        if (h.a < 1e-10) { c1 = 0.5f; c2 = 1.0f / 6.0f; }
        else {
            const float sa = s / a;
            c1 = 2.0f * sa * sa;
            if (a < 0.5f)
                c2 = 1.0f / 6.0f + a2 * (-1.0f / 120.0f + a2 * (1.0f / 5040.0f + a2 * (-1.0f / 362880.0f +
                                                                                         a2 * (1.0f / 39916800.0f))));
            else c2 = (1.0f - 2.0f * s * c / a) / a2;
        }
    } else {
        c1 = -0.5f;
        if (h.a < 1e-10) c2 = 1.0f / 12.0f;
        else if (a < 0.5f)
            c2 = 1.0f / 12.0f + a2 * (1.0f / 720.0f + a2 * (1.0f / 30240.0f + a2 * (1.0f / 1209600.0f +
                                                                                     a2 * (1.0f / 47900160.0f))));
        else c2 = 1.0f / a2 - c / (2.0f * a * s);
    }
    return TanOpF{cvt<float>(thd), c1, c2};
}

__device__ __forceinline__ M3<float> tangent_mat_f32(const TanOpF& o) {
    M3<float> S = skew(o.th);
    M3<float> S2 = mul(S, S);
    M3<float> r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.m[k] = ((k % 4 == 0) ? 1.0f : 0.0f) + o.c1 * S.m[k] + o.c2 * S2.m[k];
    return r;
}

template <class T>
__device__ __forceinline__ M3<T> tangent_op(V3<double> thd, bool inverse) {
    double a = norm(thd);
    V3<T> th = cvt<T>(thd);
    M3<T> S = skew(th);
    M3<T> S2 = mul(S, S);
    T c1, c2;
    if (!inverse) {
        if (a < 1e-10) { c1 = T(0.5); c2 = T(1.0 / 6.0); }
        else { c1 = T((1.0 - cos(a)) / (a * a)); c2 = T((1.0 - sin(a) / a) / (a * a)); }
    } else {
        c1 = T(-0.5);
        if (a < 1e-10) c2 = T(1.0 / 12.0);
        else c2 = T(1.0 / (a * a) - (1.0 + cos(a)) / (2.0 * a * sin(a)));
    }
    M3<T> r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.m[k] = ((k % 4 == 0) ? T(1) : T(0)) + c1 * S.m[k] + c2 * S2.m[k];
    return r;
}

// The symmetrised element tangent Ks = (K + K^T) / 2 in 3x3 group blocks
// (b, c) over {node-1 translations, node-1 rotations, node-2 translations,
// node-2 rotations}.  The node-2 translation rows are the negated node-1
// translation rows (translational equilibrium of the element: f_2 = -f_1 for
// the axial and shear forces, and the B-operator rows of node 2 are those of
// node 1 negated), and column group 2 is column group 0 negated (the
// [x, y, -x, z] form), so every block is +-(a block over groups {0, 1, 3}),
// possibly transposed:
//   Ks(2, c) = -Ks(0, c) for c != 2, Ks(b, 2) = -Ks(b, 0) for b != 2,
//   Ks(2, 2) = Ks(0, 0), Ks(c, b) = Ks(b, c)^T.
// Six blocks carry everything: the symmetric diagonals (0,0), (1,1), (3,3)
// (upper triangles, 6 values each) and (0,1), (0,3), (1,3) (9 values each).
template <class T>
struct KSym {
    T d0[6], d1[6], d3[6];
    T o01[9], o03[9], o13[9];
};

// Upper-triangle index of (i, j) in a symmetric 3x3 block.
__device__ __forceinline__ constexpr int sym_idx(int i, int j) {
    return i <= j ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5));
}

// Entry (i, j) of block (b, c) of Ks.
template <class T>
__device__ __forceinline__ T ksv(const KSym<T>& S, int b, int c, int i, int j) {
    const bool neg = (b == 2) != (c == 2);
    int bb = b == 2 ? 0 : b, cc = c == 2 ? 0 : c;
    if (bb > cc) {
        const int t = bb; bb = cc; cc = t;
        const int u = i; i = j; j = u;
    }
    T v;
    if (bb == cc) v = (bb == 0 ? S.d0 : (bb == 1 ? S.d1 : S.d3))[sym_idx(i, j)];
    else if (bb == 0) v = (cc == 1 ? S.o01 : S.o03)[3 * i + j];
    else v = S.o13[3 * i + j];
    return neg ? -v : v;
}

constexpr int BEAM_BLOCK = 64;

// Element e: global-frame force fe[12] (FP64), handed to on_force(fe) as soon
// as it is formed (so its registers are free during the tangent), and, when
// want_K, the symmetrised tangent S (precision TT).
template <class TT, class OnForce>
__device__ __forceinline__ void beam_core(int e, const double* __restrict__ X, const double* __restrict__ u,
                                          const int* __restrict__ conn, const BeamConst& bc, OnForce&& on_force,
                                          KSym<TT>& S, bool want_K, int* __restrict__ flags) {
    const int n1 = conn[2 * e], n2 = conn[2 * e + 1];
    V3<double> P1{X[3 * n1], X[3 * n1 + 1], X[3 * n1 + 2]};
    V3<double> P2{X[3 * n2], X[3 * n2 + 1], X[3 * n2 + 2]};
    const double* u1 = u + 6 * (size_t)n1;
    const double* u2 = u + 6 * (size_t)n2;
    V3<double> t1{u1[0], u1[1], u1[2]}, r1{u1[3], u1[4], u1[5]};
    V3<double> t2{u2[0], u2[1], u2[2]}, r2{u2[3], u2[4], u2[5]};

    // reference frame R_0 (rows e_x, e_y, e_z)
    V3<double> d0 = P2 - P1;
    constexpr bool FAST = std::is_same<TT, float>::value;   // reciprocal-multiply normalisations
    double L0 = norm(d0);
    V3<double> ex0;
    if (FAST) {
        const double i0 = 1.0 / L0;
        ex0 = V3<double>{d0.x * i0, d0.y * i0, d0.z * i0};
    } else {
        ex0 = V3<double>{d0.x / L0, d0.y / L0, d0.z / L0};
    }
    V3<double> vr;
    if (bc.has_vref) {
        vr = V3<double>{bc.vref[0], bc.vref[1], bc.vref[2]};
    } else {
        double ax = fabs(ex0.x), ay = fabs(ex0.y), az = fabs(ex0.z);
        int k = (ax <= ay && ax <= az) ? 0 : ((ay <= az) ? 1 : 2);
        vr = V3<double>{k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
    }
    V3<double> ez = cross(ex0, vr);
    double nz = norm(ez);
    nz = nz > 1e-10 ? nz : 1e-10;
    if (FAST) {
        const double iz = 1.0 / nz;
        ez = V3<double>{ez.x * iz, ez.y * iz, ez.z * iz};
    } else {
        ez = V3<double>{ez.x / nz, ez.y / nz, ez.z / nz};
    }
    V3<double> ey = cross(ez, ex0);
    M3<double> R0;
    R0.m[0] = ex0.x; R0.m[1] = ex0.y; R0.m[2] = ex0.z;
    R0.m[3] = ey.x;  R0.m[4] = ey.y;  R0.m[5] = ey.z;
    R0.m[6] = ez.x;  R0.m[7] = ez.y;  R0.m[8] = ez.z;

    // deformed chord
    V3<double> dd = (P2 + t2) - (P1 + t1);
    double Ld = norm(dd);
    V3<double> exd;
    if (FAST) {
        const double id = 1.0 / Ld;
        exd = V3<double>{dd.x * id, dd.y * id, dd.z * id};
    } else {
        exd = V3<double>{dd.x / Ld, dd.y / Ld, dd.z / Ld};
    }
    if (!(L0 >= 1e-15) || !(Ld >= 1e-15)) atomicOr(flags, FLAG_DEGENERATE);

    // minimal rotation e_x0 -> e_xd (_beam_cr.py:790-828)
    M3<double> Rrod;
    {
        V3<double> v = cross(ex0, exd);
        double s = norm(v);
        double c = dot(ex0, exd);
        if (s < 1e-14) {
            if (c > 0.0) {
                Rrod = eye3<double>();
            } else {
                double ax = fabs(ex0.x), ay = fabs(ex0.y), az = fabs(ex0.z);
                int k = (ax <= ay && ax <= az) ? 0 : ((ay <= az) ? 1 : 2);
                V3<double> perp{k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
                V3<double> axis = cross(ex0, perp);
                double an = norm(axis);
                axis = V3<double>{axis.x / an, axis.y / an, axis.z / an};
                double aa[3] = {axis.x, axis.y, axis.z};
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) Rrod.m[3 * i + j] = 2.0 * aa[i] * aa[j] - (i == j ? 1.0 : 0.0);
            }
        } else {
            M3<double> V = skew(v);
            M3<double> V2 = mul(V, V);
            double coef = (1.0 - c) / (s * s);
#pragma unroll
            for (int k = 0; k < 9; ++k) Rrod.m[k] = (((k % 4 == 0) ? 1.0 : 0.0) + V.m[k]) + coef * V2.m[k];
        }
    }
    M3<double> Rcr = mulNT(R0, Rrod);  // R_0 R_rod^T

    // nodal rotations relative to the corotated frame
    HalfAngle hr1, hr2, ht1, ht2;
    M3<double> Rd1 = mulNT(mul(Rcr, rotvec_to_rotmat(r1, &hr1)), R0);
    M3<double> Rd2 = mulNT(mul(Rcr, rotvec_to_rotmat(r2, &hr2)), R0);
    V3<double> th1 = rotmat_to_rotvec<FAST>(Rd1, &ht1);
    V3<double> th2 = rotmat_to_rotvec<FAST>(Rd2, &ht2);
    // FP32 tangent operators' coefficients, formed while the half angles are at hand
    TanOpF tos1{}, tos2{}, toi1{}, toi2{};
    if constexpr (FAST) {
        tos1 = tangent_coef_f32(r1, hr1, false);
        tos2 = tangent_coef_f32(r2, hr2, false);
        toi1 = tangent_coef_f32(th1, ht1, true);
        toi2 = tangent_coef_f32(th2, ht2, true);
    }
    double dax = Ld - L0;

    // local Timoshenko stiffness coefficients (_beam_cr.py:911-998)
    const double L = L0;
    double ea, gj, k15, k55, k511, m15, m44, m410;
    if (FAST) {
        // one reciprocal of L, L^2 and of each shear-correction denominator
        const double iL = 1.0 / L, iL2 = iL * iL;
        double phy = (12.0 * bc.E * bc.Iy / (bc.ky * bc.G * bc.A)) * iL2;
        double phz = (12.0 * bc.E * bc.Iz / (bc.kz * bc.G * bc.A)) * iL2;
        if (bc.scf > 0.0) {
            phy = phy * (1.0 / (1.0 + bc.scf * (L * L * bc.A / (12.0 * bc.Iy))));
            phz = phz * (1.0 / (1.0 + bc.scf * (L * L * bc.A / (12.0 * bc.Iz))));
        }
        const double idy = 1.0 / (1.0 + phy), idz = 1.0 / (1.0 + phz);
        ea = bc.E * bc.A * iL;
        gj = bc.G * bc.J * iL;
        k15 = 6.0 * (bc.E * bc.Iz * iL2) * idz;
        k55 = (4.0 + phz) * (bc.E * bc.Iz * iL) * idz;
        k511 = (2.0 - phz) * (bc.E * bc.Iz * iL) * idz;
        m15 = 6.0 * (bc.E * bc.Iy * iL2) * idy;
        m44 = (4.0 + phy) * (bc.E * bc.Iy * iL) * idy;
        m410 = (2.0 - phy) * (bc.E * bc.Iy * iL) * idy;
    } else {
        double phy = 12.0 * bc.E * bc.Iy / (bc.ky * bc.G * bc.A * L * L);
        double phz = 12.0 * bc.E * bc.Iz / (bc.kz * bc.G * bc.A * L * L);
        if (bc.scf > 0.0) {
            phy = phy * (1.0 / (1.0 + bc.scf * (L * L * bc.A / (12.0 * bc.Iy))));
            phz = phz * (1.0 / (1.0 + bc.scf * (L * L * bc.A / (12.0 * bc.Iz))));
        }
        const double dy = 1.0 + phy, dz = 1.0 + phz;
        const double L2 = L * L;
        ea = bc.E * bc.A / L;
        gj = bc.G * bc.J / L;
        k15 = 6.0 * (bc.E * bc.Iz / L2) / dz;          // Ke[1][5]
        k55 = (4.0 + phz) * (bc.E * bc.Iz / L) / dz;   // Ke[5][5]
        k511 = (2.0 - phz) * (bc.E * bc.Iz / L) / dz;  // Ke[5][11]
        m15 = 6.0 * (bc.E * bc.Iy / L2) / dy;          // -Ke[2][4]
        m44 = (4.0 + phy) * (bc.E * bc.Iy / L) / dy;   // Ke[4][4]
        m410 = (2.0 - phy) * (bc.E * bc.Iy / L) / dy;  // Ke[4][10]
    }

    // corotated element force f_cr = Ke d_cr (only nonzero d_cr entries used)
    double fcr[12];
    fcr[0] = -ea * dax;
    fcr[6] = ea * dax;
    fcr[3] = gj * th1.x - gj * th2.x;
    fcr[9] = -gj * th1.x + gj * th2.x;
    fcr[1] = k15 * th1.z + k15 * th2.z;
    fcr[7] = -k15 * th1.z - k15 * th2.z;
    fcr[5] = k55 * th1.z + k511 * th2.z;
    fcr[11] = k511 * th1.z + k55 * th2.z;
    fcr[2] = -m15 * th1.y - m15 * th2.y;
    fcr[8] = m15 * th1.y + m15 * th2.y;
    fcr[4] = m44 * th1.y + m410 * th2.y;
    fcr[10] = m410 * th1.y + m44 * th2.y;

    {
        double fe[12];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            V3<double> fb = mulTv(Rcr, V3<double>{fcr[3 * b], fcr[3 * b + 1], fcr[3 * b + 2]});
            fe[3 * b] = fb.x;
            fe[3 * b + 1] = fb.y;
            fe[3 * b + 2] = fb.z;
        }
        on_force(fe);
    }
    if (!want_K) return;

    // ---- analytical tangent in precision TT ----
    using T = TT;
    M3<T> Rc = cvt<T>(Rcr);
    M3<T> dpsi = cvt<T>(mul(Rcr, skew(exd)));
    {
        T il = T(1.0 / Ld);
#pragma unroll
        for (int k = 0; k < 9; ++k) dpsi.m[k] = il * dpsi.m[k];
    }
    M3<T> Ti1, Ti2, Ts1, Ts2;
    if constexpr (std::is_same<T, float>::value) {
        Ti1 = tangent_mat_f32(toi1);
        Ti2 = tangent_mat_f32(toi2);
        Ts1 = tangent_mat_f32(tos1);
        Ts2 = tangent_mat_f32(tos2);
    } else {
        Ti1 = tangent_op<T>(th1, true);
        Ti2 = tangent_op<T>(th2, true);
        Ts1 = tangent_op<T>(r1, false);
        Ts2 = tangent_op<T>(r2, false);
    }
    M3<T> P1m = mul(Ti1, dpsi);            // B[3:6, 0:3]  (B[3:6, 6:9] = -P1)
    M3<T> Q1m = mul(mul(Ti1, Rc), Ts1);    // B[3:6, 3:6]
    M3<T> P2m = mul(Ti2, dpsi);            // B[9:12, 0:3] (B[9:12, 6:9] = -P2)
    M3<T> Q2m = mul(mul(Ti2, Rc), Ts2);    // B[9:12, 9:12]
    V3<T> ex = cvt<T>(exd);

    const T tea = T(ea), tgj = T(gj), tk15 = T(k15), tk55 = T(k55), tk511 = T(k511);
    const T tm15 = T(m15), tm44 = T(m44), tm410 = T(m410);

    // (Ke B) rows as x / y / z parts; KBx[r] etc. are 3-vectors.  Rows 6-8 (node-2
    // translations) are rows 0-2 negated and are not formed.
    T KBx[12][3], KBy[12][3], KBz[12][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T exj = j == 0 ? ex.x : (j == 1 ? ex.y : ex.z);
        // axial row 0: B[6] = [-e_x, 0, e_x, 0]
        KBx[0][j] = tea * exj;  KBy[0][j] = T(0); KBz[0][j] = T(0);
        // torsion rows 3 / 9
        KBx[3][j] = tgj * P1m(0, j) - tgj * P2m(0, j);
        KBy[3][j] = tgj * Q1m(0, j);
        KBz[3][j] = -tgj * Q2m(0, j);
        KBx[9][j] = -KBx[3][j]; KBy[9][j] = -KBy[3][j]; KBz[9][j] = -KBz[3][j];
        // bending x-y plane rows 1, 5, 11 (rotation component z = row 2 of P/Q)
        KBx[1][j] = tk15 * P1m(2, j) + tk15 * P2m(2, j);
        KBy[1][j] = tk15 * Q1m(2, j);
        KBz[1][j] = tk15 * Q2m(2, j);
        KBx[5][j] = tk55 * P1m(2, j) + tk511 * P2m(2, j);
        KBy[5][j] = tk55 * Q1m(2, j);
        KBz[5][j] = tk511 * Q2m(2, j);
        KBx[11][j] = tk511 * P1m(2, j) + tk55 * P2m(2, j);
        KBy[11][j] = tk511 * Q1m(2, j);
        KBz[11][j] = tk55 * Q2m(2, j);
        // bending x-z plane rows 2, 4, 10 (rotation component y = row 1)
        KBx[2][j] = -tm15 * P1m(1, j) - tm15 * P2m(1, j);
        KBy[2][j] = -tm15 * Q1m(1, j);
        KBz[2][j] = -tm15 * Q2m(1, j);
        KBx[4][j] = tm44 * P1m(1, j) + tm410 * P2m(1, j);
        KBy[4][j] = tm44 * Q1m(1, j);
        KBz[4][j] = tm410 * Q2m(1, j);
        KBx[10][j] = tm410 * P1m(1, j) + tm44 * P2m(1, j);
        KBy[10][j] = tm410 * Q1m(1, j);
        KBz[10][j] = tm44 * Q2m(1, j);
    }

    // block row b of K: X_b = Rc^T KBx_b + Rc^T S(f_b) dpsi, Y_b = Rc^T KBy_b,
    // Z_b = Rc^T KBz_b; K(b, .) = [X_b, Y_b, -X_b, Z_b]
    auto block_row = [&](int b, M3<T>& Xo, M3<T>& Yo, M3<T>& Zo) {
        M3<T> Ax, Ay, Az;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Ax.m[3 * i + j] = KBx[3 * b + i][j];
                Ay.m[3 * i + j] = KBy[3 * b + i][j];
                Az.m[3 * i + j] = KBz[3 * b + i][j];
            }
        M3<T> W = mul(mulTN(Rc, skew(V3<T>{T(fcr[3 * b]), T(fcr[3 * b + 1]), T(fcr[3 * b + 2])})), dpsi);
        Xo = mulTN(Rc, Ax);
#pragma unroll
        for (int k = 0; k < 9; ++k) Xo.m[k] = Xo.m[k] + W.m[k];
        Yo = mulTN(Rc, Ay);
        Zo = mulTN(Rc, Az);
    };
    // Ks(b, c)(i, j) = (K(b, c)(i, j) + K(c, b)(j, i)) / 2 over groups {0, 1, 3},
    // each block folded as soon as both of its sources exist
    M3<T> X0, Y0, Z0, X1, Y1, Z1, X3, Y3, Z3;
    block_row(0, X0, Y0, Z0);
    block_row(1, X1, Y1, Z1);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            S.d0[sym_idx(i, j)] = T(0.5) * (X0(i, j) + X0(j, i));
            S.d1[sym_idx(i, j)] = T(0.5) * (Y1(i, j) + Y1(j, i));
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) S.o01[3 * i + j] = T(0.5) * (Y0(i, j) + X1(j, i));
    block_row(3, X3, Y3, Z3);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S.o03[3 * i + j] = T(0.5) * (Z0(i, j) + X3(j, i));
            S.o13[3 * i + j] = T(0.5) * (Z1(i, j) + Y3(j, i));
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) S.d3[sym_idx(i, j)] = T(0.5) * (Z3(i, j) + Z3(j, i));
}

// Staged path: per-element records (any connectivity).
template <class TT>
__global__ void __launch_bounds__(BEAM_BLOCK)
beam_kernel(int n_elems, const double* __restrict__ X, const double* __restrict__ u,
            const int* __restrict__ conn, BeamConst bc, double* __restrict__ fe_stage,
            TT* __restrict__ Ke_stage, int* __restrict__ flags, const double* __restrict__ stop) {
    if (!nc_go(stop)) return;   // the Newton call this attempt belongs to has ended
    // tail lanes recompute the last element (nothing of theirs is stored) so
    // that every lane reaches the LDS barriers of the tangent write-out
    const int e_lane = blockIdx.x * BEAM_BLOCK + threadIdx.x;
    const bool live = e_lane < n_elems;
    const int e = live ? e_lane : n_elems - 1;
    KSym<TT> S;
    beam_core<TT>(e, X, u, conn, bc, [&](const double* fe) {
        if (fe_stage && live) {
#pragma unroll
            for (int k = 0; k < 12; ++k) fe_stage[12 * (size_t)e + k] = fe[k];
        }
    }, S, Ke_stage != nullptr, flags);
    if (!Ke_stage) return;  // uniform across the launch
    // block row b of the 144-value record = one 36-value chunk, through LDS
    __shared__ TT tile[BEAM_BLOCK * 36];
    const int e0 = blockIdx.x * BEAM_BLOCK;
    const int cnt = min(BEAM_BLOCK, n_elems - e0);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) tile[threadIdx.x * 36 + c * 9 + 3 * i + j] = ksv(S, b, c, i, j);
        __syncthreads();
        TT* dst = Ke_stage + (size_t)e0 * 144 + 36 * b;
        for (int k = threadIdx.x; k < cnt * 36; k += BEAM_BLOCK) {
            const int el = k / 36, j = k - 36 * (k / 36);
            dst[(size_t)el * 144 + j] = tile[k];
        }
        __syncthreads();
    }
}

// Direct path (chains with consecutive node numbers: element e joins n and
// n + 1): eflag bit0 = node 1 shared with element e-1, bit1 = node 2 shared
// with element e+1 (each such node has exactly these two elements).
// eslot[16 e + lb] = global block index of group block lb.
//
// The tangent rows of a node n of such a chain hold its beam entries as one
// contiguous run of the scalar CSR row: column groups 2(n-1) .. 2(n+1)+1, 18
// entries (12 at a chain end).  A workgroup builds the runs of the rows of
// its owned nodes in LDS, IMG_ROWS scalar rows per lane at a time: each lane
// writes its element's node-2 rows, then adds its node-1 rows into the previous
// lane's image (element order, so the sums are fixed), and the workgroup stores
// the runs with consecutive lanes on consecutive addresses.  The first node of
// a chain is written by its element directly.
//
// Two waves per SIMD: the FP32 form needs about 170 registers, the FP64 form
// is held to 256 (a few values spill to scratch; at its natural 340 it would
// run one wave per SIMD and take 97 instead of 81 us on 16 x 91 wires).
constexpr int DIRECT_OWN = BEAM_BLOCK - 1;
constexpr int RUN = 18;       // entries of an interior row run

template <class TT>
__global__ void __launch_bounds__(BEAM_BLOCK, 2)
beam_direct_kernel(int n_elems, const double* __restrict__ X, const double* __restrict__ u,
                   const int* __restrict__ conn, BeamConst bc, const int8_t* __restrict__ eflag,
                   const int* __restrict__ eslot, const int* __restrict__ gptr, double mass_coef,
                   const double* __restrict__ M, double* __restrict__ fint, TT* __restrict__ K,
                   int* __restrict__ flags, const double* __restrict__ stop) {
    if (!nc_go(stop)) return;   // the Newton call this attempt belongs to has ended
    // scalar rows per lane in one LDS image: a whole row group (13.8 KB) in
    // FP32, one row (9.2 KB) in FP64, where three would take 29 KB and leave
    // room for one wave per SIMD
    constexpr int IMG_ROWS = sizeof(TT) == 4 ? 3 : 1;
    // node 2's mass terms: loaded before the element math in FP32, after it in
    // FP64, whose registers are full during the math
    constexpr bool MASS_EARLY = sizeof(TT) == 4;
    const int l = threadIdx.x;
    const int e_lane = blockIdx.x * DIRECT_OWN + l;
    const bool live = e_lane < n_elems;
    const int e = live ? e_lane : n_elems - 1;
    const int fl = eflag[e];
    const bool prv = live && (fl & 1), nxt = (fl & 2) != 0;
    const bool own = live && l < DIRECT_OWN;
    const int n1 = conn[2 * e], n2 = conn[2 * e + 1];
    // Row offsets and block slots of the rows this lane writes, loaded before
    // the element math so that their latency hides behind it (loaded at the
    // write-out they cost one dependent round trip per row).  Index 0/1 of
    // gq/bq: node 2's rows, then the chain-start rows of node 1 (whose mass
    // terms are loaded where they are used: one lane in 150).
    int gq[2][3] = {}, bq[2][2] = {};
    double mq[6] = {};
    const bool first = own && !prv;
    if (K) {   // uniform
#pragma unroll
        for (int k = 0; k < 3; ++k) gq[0][k] = gptr[2 * n2 + k];
        bq[0][0] = eslot[16 * (size_t)e + 8];
        bq[0][1] = eslot[16 * (size_t)e + 12];
        if (MASS_EARLY && mass_coef != 0.0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) mq[k] = M[6 * (size_t)n2 + k];
        }
        if (first) {
#pragma unroll
            for (int k = 0; k < 3; ++k) gq[1][k] = gptr[2 * n1 + k];
            bq[1][0] = eslot[16 * (size_t)e];
            bq[1][1] = eslot[16 * (size_t)e + 4];
        }
    }
    // LDS: the node-1 force parts first, then the row image in the same bytes
    constexpr int IMG_BYTES = BEAM_BLOCK * IMG_ROWS * RUN * (int)sizeof(TT);
    constexpr int F_BYTES = BEAM_BLOCK * 6 * (int)sizeof(double);
    __shared__ __align__(16) unsigned char smem[IMG_BYTES > F_BYTES ? IMG_BYTES : F_BYTES];
    __shared__ int2 dst[BEAM_BLOCK * IMG_ROWS];   // per image row: CSR offset (-1: not owned), run length
    double* fsh = reinterpret_cast<double*>(smem);
    KSym<TT> S;
    beam_core<TT>(e, X, u, conn, bc, [&](const double* fe) {
        // node-1 force part for the previous lane
#pragma unroll
        for (int d = 0; d < 6; ++d) fsh[l * 6 + d] = fe[d];
        __syncthreads();
        if (fint && own) {
#pragma unroll
            for (int d = 0; d < 6; ++d) fint[6 * (size_t)n2 + d] = nxt ? fe[6 + d] + fsh[(l + 1) * 6 + d] : fe[6 + d];
            if (!prv) {
#pragma unroll
                for (int d = 0; d < 6; ++d) fint[6 * (size_t)n1 + d] = fe[d];
            }
        }
    }, S, K != nullptr, flags);
    if (!K) return;  // uniform across the launch
    __syncthreads();   // every lane has read the force parts before the image overwrites them

    // lumped-mass terms of node 2's diagonal entries
    TT mt[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        mt[k] = mass_coef != 0.0 ? TT(mass_coef * (MASS_EARLY ? mq[k] : M[6 * (size_t)n2 + k])) : TT(0);
    // first node of a chain: its rows hold this element's node-1 blocks only
    if (first) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int r0 = gq[1][p], deg = gq[1][p + 1] - r0;
            const int b0 = bq[1][p];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const size_t d = 9 * (size_t)r0 + (size_t)i * 3 * deg + 3 * (size_t)(b0 - r0);
                const TT m = mass_coef != 0.0 ? TT(mass_coef * M[6 * (size_t)n1 + 3 * p + i]) : TT(0);
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        TT v = ksv(S, p, c, i, j);
                        if (c == p && j == i) v += m;
                        K[d + 3 * c + j] = v;
                    }
            }
        }
    }
    TT* img = reinterpret_cast<TT*>(smem);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int r0 = gq[0][p], deg = gq[0][p + 1] - r0;
        const int b0 = bq[0][p];
#pragma unroll
        for (int i0 = 0; i0 < 3; i0 += IMG_ROWS) {
            // own node-2 rows: columns [node 1 | node 2 | next element's node 2]
#pragma unroll
            for (int ii = 0; ii < IMG_ROWS; ++ii) {
                const int i = i0 + ii;
                TT* row = img + (l * IMG_ROWS + ii) * RUN;
                const TT m = mt[3 * p + i];
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        TT v = ksv(S, 2 + p, c, i, j);
                        if (c == 2 + p && j == i) v += m;
                        row[3 * c + j] = v;
                    }
#pragma unroll
                for (int k = 12; k < RUN; ++k) row[k] = TT(0);
                dst[l * IMG_ROWS + ii] = int2{own ? 9 * r0 + i * 3 * deg + 3 * (b0 - r0) : -1, nxt ? RUN : 12};
            }
            __syncthreads();
            // node-1 rows into the previous lane's node-2 rows (same node)
            if (prv && l > 0) {
#pragma unroll
                for (int ii = 0; ii < IMG_ROWS; ++ii) {
                    TT* row = img + ((l - 1) * IMG_ROWS + ii) * RUN + 6;
#pragma unroll
                    for (int c = 0; c < 4; ++c)
#pragma unroll
                        for (int j = 0; j < 3; ++j) row[3 * c + j] += ksv(S, p, c, i0 + ii, j);
                }
            }
            __syncthreads();
            // entry t = l + 64 it of the image: row rr = t / RUN, column k = t % RUN.
            // The image holds BEAM_BLOCK * IMG_ROWS rows, so every t < BEAM_BLOCK *
            // IMG_ROWS * RUN reads inside it; rows of the last lane (not owned)
            // carry offset -1.  Each group of U entries reads its offsets and
            // values first and stores after, so the LDS reads of the group
            // overlap instead of each store waiting on its own chain of reads.
            {
                constexpr int NIT = IMG_ROWS * RUN;   // BEAM_BLOCK * IMG_ROWS * RUN / BEAM_BLOCK
                constexpr int U = 6;
                static_assert(NIT % U == 0, "store loop unroll");
                constexpr int STEP_R = BEAM_BLOCK / RUN, STEP_K = BEAM_BLOCK % RUN;
                int rr = l / RUN, k = l % RUN, t = l;
#pragma unroll 1
                for (int it0 = 0; it0 < NIT; it0 += U) {
                    int2 dd[U];
                    TT vv[U];
                    int kk[U];
#pragma unroll
                    for (int j = 0; j < U; ++j) {
                        dd[j] = dst[rr];
                        vv[j] = img[t];
                        kk[j] = k;
                        t += BEAM_BLOCK;
                        rr += STEP_R;
                        k += STEP_K;
                        if (k >= RUN) { k -= RUN; ++rr; }
                    }
#pragma unroll
                    for (int j = 0; j < U; ++j)
                        if ((dd[j].x >= 0) & (kk[j] < dd[j].y)) K[(size_t)dd[j].x + kk[j]] = vv[j];
                }
            }
            __syncthreads();
        }
    }

}

}  // namespace xk

// host-side launchers
namespace xk {
hipError_t launch_beam(int prec, int n_elems, const double* X, const double* u, const int* conn,
                       const BeamConst& bc, double* fe_stage, void* Ke_stage, int* flags,
                       hipStream_t st, const double* stop) {
    if (n_elems <= 0) return hipSuccess;
    dim3 blk(BEAM_BLOCK), grd((n_elems + BEAM_BLOCK - 1) / BEAM_BLOCK);
    if (prec == 0)
        hipLaunchKernelGGL(beam_kernel<double>, grd, blk, 0, st, n_elems, X, u, conn, bc, fe_stage,
                           (double*)Ke_stage, flags, stop);
    else
        hipLaunchKernelGGL(beam_kernel<float>, grd, blk, 0, st, n_elems, X, u, conn, bc, fe_stage,
                           (float*)Ke_stage, flags, stop);
    return hipGetLastError();
}

hipError_t launch_beam_direct(int prec, int n_elems, const double* X, const double* u, const int* conn,
                              const BeamConst& bc, const int8_t* eflag, const int* eslot, const int* gptr,
                              double mass_coef, const double* M, double* f_int, void* K, int* flags,
                              hipStream_t st, const double* stop) {
    if (n_elems <= 0) return hipSuccess;
    dim3 blk(BEAM_BLOCK), grd((n_elems + DIRECT_OWN - 1) / DIRECT_OWN);
    if (prec == 0)
        hipLaunchKernelGGL(beam_direct_kernel<double>, grd, blk, 0, st, n_elems, X, u, conn, bc, eflag, eslot,
                           gptr, mass_coef, M, f_int, (double*)K, flags, stop);
    else
        hipLaunchKernelGGL(beam_direct_kernel<float>, grd, blk, 0, st, n_elems, X, u, conn, bc, eflag, eslot,
                           gptr, mass_coef, M, f_int, (float*)K, flags, stop);
    return hipGetLastError();
}
}  // namespace xk
