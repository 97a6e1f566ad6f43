// Beam-beam contact pairs: one pair per lane.
//
// Covers, per pair and in this order (all selectable by mode bits):
//   narrowphase   clamped segment-segment closest point
//                 (xkep_cae/contact/geometry/_compute.py:55-151)
//   frame         parallel transport of t1 along the minimal rotation
//                 n_old -> n_new, Gram-Schmidt and axis fallbacks (:154-261)
//   hysteresis    INACTIVE -> ACTIVE at gap <= g_on, ACTIVE/SLIDING ->
//                 INACTIVE at gap >= g_off (contact/_manager_process.py:475-494)
//   Huber         p_n = huber(k_pen (-gap), k_pen / smoothing_delta) for
//                 non-INACTIVE pairs (contact_force/strategy.py:104-170)
//   friction      Coulomb return mapping for every pair with p_n > 0,
//                 k_t initialised on first use, z_t <- q, status <- ACTIVE /
//                 SLIDING (friction/_assembly.py:260-342, law_friction.py:89-153)
//   tangents      Huber material + geometric, friction material
//                 sum_ab D_ab t_a t_b^T and geometric M / dist
//                 (contact_force/strategy.py:172-254, friction/_assembly.py:108-257)
//   coating       with the Kelvin-Voigt coating on (pp.coat_on): core gap and
//                 compression cc = max(0, coat_total - gap_core), activation by
//                 cc > 0 (_manager_process.py:351-360, :486-494); by mode bit
//                 the force k cc + c (cc - cc_prev) / dt along -n (COAT_F), the
//                 coating Coulomb return mapping (k_t = k k_t_ratio, p_n = k cc;
//                 COAT_FRIC), the stiffness k_eff n n^T (COAT_K) and T D T^T
//                 (COAT_KF) (coating/strategy.py:99-370)
//   probe         (PAIR_PROBE, alone) the compression at x only: cneed[i] =
//                 cc > 0; nothing else is written
//
// Every 3x3 node block (k, l) of a pair's contribution to K is c_k c_l G with
// one 3x3 G per pair and c = [(1-s), s, -(1-t), -t]; every node force is
// c_k F with one 3-vector F.  The coating terms have this form too: their
// weights [-(1-s), -s, 1-t, t] are -c, so they enter F with a minus sign and
// G as they are (w_k w_l = c_k c_l).  The kernel therefore stages per contributing
// pair only a force record c[4], F[3] (FP64) and a tangent record G[9], c[4]
// (FP64 or FP32), at the pair's contributing slot of the increment; the
// gather kernel expands them.  Pair blocks are mapped XCD-contiguously, so
// the node coordinates of nearby pairs are fetched into one L2.
//
// This translation unit is compiled with FP contraction off so the
// branch-deciding quantities (parallel test, clamps, gaps, hysteresis
// thresholds) are formed with the same roundings as the NumPy reference.

#include "xk_common.h"
#include "xk_structs.h"

namespace xk {

__device__ __forceinline__ double clip01(double x) { return fmin(fmax(x, 0.0), 1.0); }
__device__ __forceinline__ float clip01f(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ V3<double> node_x(const double* __restrict__ Xr, const double* __restrict__ u, int k) {
    V3<double> x{Xr[3 * k], Xr[3 * k + 1], Xr[3 * k + 2]};
    if (u) x = x + V3<double>{u[6 * (size_t)k], u[6 * (size_t)k + 1], u[6 * (size_t)k + 2]};
    return x;
}

__device__ __forceinline__ V3<double> node_du(const double* __restrict__ u, const double* __restrict__ ur, int k) {
    const size_t o = 6 * (size_t)k;
    V3<double> a{u[o], u[o + 1], u[o + 2]};
    if (ur) a = a - V3<double>{ur[o], ur[o + 1], ur[o + 2]};
    return a;
}

template <class T>
__device__ __forceinline__ V3<T> min_axis(V3<T> v) {
    T ax = fabs(v.x), ay = fabs(v.y), az = fabs(v.z);
    int k = (ax <= ay && ax <= az) ? 0 : ((ay <= az) ? 1 : 2);
    return V3<T>{T(k == 0 ? 1 : 0), T(k == 1 ? 1 : 0), T(k == 2 ? 1 : 0)};
}

// Clamped segment-segment closest points (_compute.py:55-151): s, t, the
// distance and the unit normal (pA - pB) / dist, (0, 0, 1) below 1e-30.
__device__ __forceinline__ void closest_points(const V3<double>& xA0, const V3<double>& xA1, const V3<double>& xB0,
                                               const V3<double>& xB1, double& s, double& t, double& dist,
                                               V3<double>& nrm) {
    V3<double> dA = xA1 - xA0, dB = xB1 - xB0, w0 = xA0 - xB0;
    double a = dot(dA, dA), b = dot(dA, dB), c = dot(dB, dB), d = dot(dA, w0), e = dot(dB, w0);
    double det = a * c - b * b;
    double ac = a * c;
    bool par = det < 1e-10 * (ac > 1e-30 ? ac : 1e-30);
    double s_unc = par ? 0.0 : (b * e - c * d) / det;
    double t_unc = par ? 0.0 : (a * e - b * d) / det;
    bool c_ok = c > 1e-10, a_ok = a > 1e-10;
    double sc = c_ok ? c : 1.0, sa = a_ok ? a : 1.0;
    s = par ? 0.0 : clip01(s_unc);
    t = par ? (c_ok ? clip01(e / sc) : 0.0) : clip01(t_unc);
    bool s_cl = (s_unc < 0.0) || (s_unc > 1.0);
    bool t_cl = (t_unc < 0.0) || (t_unc > 1.0);
    if (s_cl && !par) t = c_ok ? clip01((b * s + e) / sc) : 0.0;
    if (t_cl && !par) {
        s = a_ok ? clip01((b * t - d) / sa) : 0.0;
        t = c_ok ? clip01((b * s + e) / sc) : 0.0;
    }
    V3<double> diff = (xA0 + s * dA) - (xB0 + t * dB);
    dist = sqrt(dot(diff, diff));
    double sd_ = dist > 1e-30 ? dist : 1e-30;
    nrm = V3<double>{diff.x / sd_, diff.y / sd_, diff.z / sd_};
    if (dist < 1e-30) nrm = V3<double>{0.0, 0.0, 1.0};
}

// Contact frame (_compute.py:154-261): n = nrm normalised, t1 = the previous
// t1 parallel-transported along the minimal rotation pn -> n, projected and
// normalised; else the previous t1 projected; else the min-|n| axis projected.
// t2 = n x t1 is never stored.  T = float only in the FP32 pair form below.
template <class T>
__device__ __forceinline__ void transport_frame(const V3<T>& nrm, const V3<T>& pn, const V3<T>& pt1, V3<T>& n,
                                                V3<T>& t1) {
    T nn_ = sqrt(dot(nrm, nrm));
    nn_ = nn_ > T(1e-30) ? nn_ : T(1e-30);
    n = V3<T>{nrm.x / nn_, nrm.y / nn_, nrm.z / nn_};
    bool has_t = sqrt(dot(pt1, pt1)) > T(1e-10);
    bool has_n = sqrt(dot(pn, pn)) > T(1e-10);
    bool done = false;
    if (has_t && has_n) {
        T on = sqrt(dot(pn, pn));
        on = on > T(1e-30) ? on : T(1e-30);
        V3<T> no{pn.x / on, pn.y / on, pn.z / on};
        V3<T> v = cross(no, n);
        T sv = sqrt(dot(v, v));
        T cv = dot(no, n);
        V3<T> tt = pt1;
        if (sv >= T(1e-12)) {
            V3<T> vxt = cross(v, pt1);
            V3<T> vxvxt = cross(v, vxt);
            T fac = (T(1) - cv) / (sv * sv);
            tt = (pt1 + vxt) + fac * vxvxt;
        }
        V3<T> raw = tt - dot(tt, n) * n;
        T rn = sqrt(dot(raw, raw));
        if (rn > T(1e-10)) {
            t1 = V3<T>{raw.x / rn, raw.y / rn, raw.z / rn};
            done = true;
        }
    }
    if (!done && has_t) {
        V3<T> raw = pt1 - dot(pt1, n) * n;
        T rn = sqrt(dot(raw, raw));
        if (rn > T(1e-10)) {
            t1 = V3<T>{raw.x / rn, raw.y / rn, raw.z / rn};
            done = true;
        }
    }
    if (!done) {
        V3<T> ref = min_axis(n);
        V3<T> raw = ref - dot(ref, n) * n;
        T rn = sqrt(dot(raw, raw));
        rn = rn > T(1e-30) ? rn : T(1e-30);
        t1 = V3<T>{raw.x / rn, raw.y / rn, raw.z / rn};
    }
}

// FP32 form of the narrowphase and frame of a pair that contributes nothing
// (status INACTIVE, no coating).  The segment differences are formed in FP64
// from the FP64 positions and rounded once; the closest point, distance and
// frame transport then run in FP32, and s, t, gap, n, t1 are stored widened.
// When `decide` (the hysteresis pass), the FP32 gap must show the pair stays
// INACTIVE: kept only if gap > g_on + margin, where the margin bounds how far
// the FP32 closest point can sit above the true minimum distance (the FP32
// parameters move by about 2^-23 ac / det, the distance by at most the segment
// lengths times that).  Nearly parallel or degenerate segments (det < 1e-4 ac,
// a or c below 1e-10), whose FP32 parameters are ill-conditioned, are never
// kept.  Returns false when the caller must evaluate the pair in FP64.
__device__ __forceinline__ bool geom_f32(const V3<double>& xA0, const V3<double>& xA1, const V3<double>& xB0,
                                         const V3<double>& xB1, double rsum, bool decide, double g_on,
                                         const V3<double>& pn, const V3<double>& pt1, double& s_o, double& t_o,
                                         double& gap_o, V3<double>& n_o, V3<double>& t1_o) {
    auto f32 = [](const V3<double>& v) { return V3<float>{(float)v.x, (float)v.y, (float)v.z}; };
    const V3<float> dA = f32(xA1 - xA0), dB = f32(xB1 - xB0), w0 = f32(xA0 - xB0);
    const float a = dot(dA, dA), b = dot(dA, dB), c = dot(dB, dB), d = dot(dA, w0), e = dot(dB, w0);
    const float ac = a * c, det = ac - b * b;
    if (!(a > 1e-10f && c > 1e-10f && det >= 1e-4f * ac)) return false;
    const float s_unc = (b * e - c * d) / det, t_unc = (a * e - b * d) / det;
    float s = clip01f(s_unc), t = clip01f(t_unc);
    if (s_unc < 0.0f || s_unc > 1.0f) t = clip01f((b * s + e) / c);
    if (t_unc < 0.0f || t_unc > 1.0f) {
        s = clip01f((b * t - d) / a);
        t = clip01f((b * s + e) / c);
    }
    const V3<float> diff = (w0 + s * dA) - t * dB;
    const float dist = sqrtf(dot(diff, diff));
    const double gap = (double)dist - rsum;
    if (decide) {
        const double margin = 2.4e-6 * ((double)sqrtf(a) + (double)sqrtf(c)) * ((double)ac / (double)det) +
                              1e-6 * (double)dist;
        if (!(gap > g_on + margin)) return false;
    }
    const float sd = dist > 1e-30f ? dist : 1e-30f;
    V3<float> nrm{diff.x / sd, diff.y / sd, diff.z / sd};
    if (dist < 1e-30f) nrm = V3<float>{0.0f, 0.0f, 1.0f};
    V3<float> n, t1;
    transport_frame(nrm, f32(pn), f32(pt1), n, t1);
    s_o = s;
    t_o = t;
    gap_o = gap;
    n_o = V3<double>{n.x, n.y, n.z};
    t1_o = V3<double>{t1.x, t1.y, t1.z};
    return true;
}

// G32: the FP32 pair form (geom_f32 for pairs that contribute nothing).
// COAT: the Kelvin-Voigt coating terms are compiled in (pp.coat_on); the
// plain instantiation carries none of their registers.
template <class GT, int BS, bool COAT, bool G32>
__global__ void __launch_bounds__(BS)
pair_kernel(int64_t n_pairs, PairSoA P, const double* __restrict__ Xr, const double* __restrict__ u,
            const double* __restrict__ u_ref, int modes, PairParams pp, double* __restrict__ nf,
            GT* __restrict__ sg, int* __restrict__ act_part, const int* __restrict__ idx) {
    // the Newton call this attempt belongs to has ended (or waits on the host):
    // the friction history must not move
    if (!nc_go(pp.stop)) return;
    const int lb = xcd_logical_block();
    const int n_blk = (int)((n_pairs + BS - 1) / BS);
    if (lb >= n_blk) return;  // padding workgroup (uniform over the workgroup)
    const int64_t w = (int64_t)lb * BS + threadIdx.x;
    int active_here = 0;
    if (w < n_pairs) {
        const int64_t i = idx ? (int64_t)idx[w] : w;
        const int64_t cap = P.cap;
        // The per-pair state streams through once per assembly (about 200 bytes a
        // pair); loaded and stored non-temporally it does not evict the node
        // vectors and records that the gather kernel reads next.
        int8_t st = __builtin_nontemporal_load(&P.status[i]);
        double s, t, gap;
        V3<double> n, t1;
        const int a0 = __builtin_nontemporal_load(&P.na0[i]), a1 = __builtin_nontemporal_load(&P.na1[i]), b0 = __builtin_nontemporal_load(&P.nb0[i]), b1 = __builtin_nontemporal_load(&P.nb1[i]);
        const double ra = __builtin_nontemporal_load(&P.ra[i]), rb = __builtin_nontemporal_load(&P.rb[i]);
        constexpr bool coat = COAT;
        double cc = 0.0;   // coating compression
        // The previous frame and the record slot are loaded with the ids, before
        // the dependent node gathers, so the three loads overlap.
        const int si = __builtin_nontemporal_load(&P.sinc[i]);
        const bool geom = (modes & PAIR_GEOM) && !(modes & PAIR_PROBE);
        V3<double> pt1{0.0, 0.0, 0.0}, pn{0.0, 0.0, 0.0};
        if (geom) {
            pt1 = V3<double>{__builtin_nontemporal_load(&P.t1[i]), __builtin_nontemporal_load(&P.t1[cap + i]), __builtin_nontemporal_load(&P.t1[2 * cap + i])};
            pn = V3<double>{__builtin_nontemporal_load(&P.n[i]), __builtin_nontemporal_load(&P.n[cap + i]), __builtin_nontemporal_load(&P.n[2 * cap + i])};
        }
        if (modes & PAIR_GEOM) {
            V3<double> xA0 = node_x(Xr, u, a0), xA1 = node_x(Xr, u, a1);
            V3<double> xB0 = node_x(Xr, u, b0), xB1 = node_x(Xr, u, b1);
            double dist;
            V3<double> nrm;
            // COAT and G32 are never both set (launch_pairs_bs)
            const bool in_f32 = G32 && st == ST_INACTIVE && geom &&
                                geom_f32(xA0, xA1, xB0, xB1, ra + rb, (modes & PAIR_HYST) != 0, pp.g_on, pn, pt1, s, t,
                                         gap, n, t1);
            if (!in_f32) closest_points(xA0, xA1, xB0, xB1, s, t, dist, nrm);
            if (coat) {
                const double cra = P.cra[i], crb = P.crb[i];
                gap = dist - (cra + crb);
                const double x = ((ra - cra) + (rb - crb)) - gap;
                cc = x > 0.0 ? x : 0.0;
            } else if (!in_f32) {
                gap = dist - (ra + rb);
            }
            if (modes & PAIR_PROBE) {
                P.cneed[i] = cc > 0.0 ? 1 : 0;
                goto done_pair;
            }

            if (!in_f32) transport_frame(nrm, pn, pt1, n, t1);
            __builtin_nontemporal_store(s, &P.s[i]);
            __builtin_nontemporal_store(t, &P.t[i]);
            __builtin_nontemporal_store(gap, &P.gap[i]);
            __builtin_nontemporal_store(n.x, &P.n[i]);
            __builtin_nontemporal_store(n.y, &P.n[cap + i]);
            __builtin_nontemporal_store(n.z, &P.n[2 * cap + i]);
            __builtin_nontemporal_store(t1.x, &P.t1[i]);
            __builtin_nontemporal_store(t1.y, &P.t1[cap + i]);
            __builtin_nontemporal_store(t1.z, &P.t1[2 * cap + i]);
            if (coat) P.cc[i] = cc;
            if (modes & PAIR_HYST) {
                const bool coat_active = coat && cc > 0.0;
                int8_t ns = st;
                if (st == ST_INACTIVE) {
                    if (gap <= pp.g_on || coat_active) ns = ST_ACTIVE;
                } else if (pp.allow_deact && gap >= pp.g_off && !coat_active) {
                    ns = ST_INACTIVE;
                }
                if (ns != st) { P.status[i] = ns; st = ns; }
            }
        } else {
            s = P.s[i]; t = P.t[i]; gap = P.gap[i];
            n = V3<double>{P.n[i], P.n[cap + i], P.n[2 * cap + i]};
            t1 = V3<double>{P.t1[i], P.t1[cap + i], P.t1[2 * cap + i]};
            if (coat) cc = P.cc[i];
        }

        // coating: history reset of a separated pair (every pair, strategy.py:220-233);
        // a pair with a coating contribution but no slot of the pattern is reported
        const bool coat_f = coat && (modes & PAIR_COAT_F);
        const bool coat_fr = coat && (modes & PAIR_COAT_FRIC);
        if (coat_fr && pp.coat_mu > 0.0 && cc <= 0.0) {
            P.cz[i] = 0.0; P.cz[cap + i] = 0.0;
            P.cstick[i] = 1;
            P.cqtn[i] = 0.0;
            P.cdiss[i] = 0.0;
        }
        if (coat && si < 0 && (modes & PAIR_COAT) && (cc > 0.0 || P.ccp[i] > 0.0))
            atomicOr(pp.flags, FLAG_COAT_UNPLACED);
        if (si >= 0 && (modes & (PAIR_HUBER | PAIR_FRIC | PAIR_TAN_C | PAIR_TAN_F | PAIR_COAT))) {
            const V3<double> t2 = cross(n, t1);
            double p_n = P.p_n[i];
            V3<double> F{0.0, 0.0, 0.0};
            const double x_pen = pp.k_pen * (-gap);
            const double dh = pp.delta_h;
            if ((modes & PAIR_HUBER) && st != ST_INACTIVE) {
                if (dh <= 0.0) p_n = x_pen > 0.0 ? x_pen : 0.0;
                else if (x_pen < -dh) p_n = 0.0;
                else if (x_pen > dh) p_n = x_pen;
                else p_n = ((x_pen + dh) * (x_pen + dh)) / (4.0 * dh);
                P.p_n[i] = p_n;
                if (p_n > 1e-30) F = p_n * n;
            }
            const double mu = pp.mu_elem ? pp.mu_eff * pp.mu_elem[pp.ea[i]] : pp.mu_eff;
            if ((modes & PAIR_FRIC) && p_n > 0.0 && mu > 0.0) {
                double kp = P.k_pen[i], kt = P.k_t[i];
                if (kp <= 0.0) {
                    kp = pp.k_pen;
                    kt = pp.k_pen * pp.k_t_ratio;
                    P.k_pen[i] = kp;
                    P.k_t[i] = kt;
                }
                V3<double> duA = (1.0 - s) * node_du(u, u_ref, a0) + s * node_du(u, u_ref, a1);
                V3<double> duB = (1.0 - t) * node_du(u, u_ref, b0) + t * node_du(u, u_ref, b1);
                V3<double> dur = duB - duA;
                double du1 = dot(dur, t1), du2 = dot(dur, t2);
                double q1 = P.z[i] + kt * du1;
                double q2 = P.z[cap + i] + kt * du2;
                double qtn = sqrt(q1 * q1 + q2 * q2);
                double fy = mu * p_n;
                bool stick = qtn <= fy;
                if (!stick) {
                    q1 = (fy * q1) / qtn;
                    q2 = (fy * q2) / qtn;
                }
                P.z[i] = q1; P.z[cap + i] = q2;
                P.stick[i] = stick ? 1 : 0;
                P.qtn[i] = qtn;
                P.diss[i] = q1 * du1 + q2 * du2;
                st = stick ? ST_ACTIVE : ST_SLIDING;
                P.status[i] = st;
                if (sqrt(q1 * q1 + q2 * q2) >= 1e-30) {
                    double f1 = fabs(q1) < 1e-30 ? 0.0 : q1;
                    double f2 = fabs(q2) < 1e-30 ? 0.0 : q2;
                    F = F + (f1 * t1 + f2 * t2);
                }
            }
            if (coat_f) {
                // Kelvin-Voigt normal force (strategy.py:99-144): node forces w_k f_n n = c_k (-f_n n)
                const double ccp = P.ccp[i];
                if (!(cc <= 0.0 && ccp <= 0.0)) {
                    double f_n = pp.coat_k * cc;
                    if (pp.coat_c > 0.0 && pp.coat_dt > 0.0) f_n = f_n + pp.coat_c * ((cc - ccp) / pp.coat_dt);
                    if (!(fabs(f_n) < 1e-30)) F = F - f_n * n;
                }
            }
            if (coat_fr) {
                // coating friction (strategy.py:194-286): return mapping of the coating history
                if (pp.coat_mu > 0.0 && cc > 0.0) {
                    const double pnc = pp.coat_k * cc, kt = pp.coat_kt;
                    V3<double> duA = (1.0 - s) * node_du(u, u_ref, a0) + s * node_du(u, u_ref, a1);
                    V3<double> duB = (1.0 - t) * node_du(u, u_ref, b0) + t * node_du(u, u_ref, b1);
                    V3<double> dur = duB - duA;
                    double du1 = dot(dur, t1), du2 = dot(dur, t2);
                    double q1 = P.cz[i] + kt * du1;
                    double q2 = P.cz[cap + i] + kt * du2;
                    double qtn = sqrt(q1 * q1 + q2 * q2);
                    double fy = pp.coat_mu * pnc;
                    bool stick = qtn <= fy;
                    if (!stick) {
                        q1 = (fy * q1) / qtn;
                        q2 = (fy * q2) / qtn;
                    }
                    P.cz[i] = q1; P.cz[cap + i] = q2;
                    P.cstick[i] = stick ? 1 : 0;
                    P.cqtn[i] = qtn;
                    P.cdiss[i] = q1 * du1 + q2 * du2;
                    F = F - (q1 * t1 + q2 * t2);
                }
            }
            if (modes & (PAIR_HUBER | PAIR_FRIC | PAIR_COAT_F | PAIR_COAT_FRIC)) {
                // c_k F at node k, where the gather's node lists read it
                const double ck[4] = {1.0 - s, s, -(1.0 - t), -t};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    double* o = nf + 3 * (size_t)pp.nqpos[4 * (size_t)si + k];
                    o[0] = ck[k] * F.x;
                    o[1] = ck[k] * F.y;
                    o[2] = ck[k] * F.z;
                }
            }
            if (modes & (PAIR_TAN_C | PAIR_TAN_F | PAIR_COAT_K | PAIR_COAT_KF)) {
                double G[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) G[k] = 0.0;
                const double nv[3] = {n.x, n.y, n.z};
                const double dist = (gap + ra) + rb;
                if ((modes & PAIR_TAN_C) && st != ST_INACTIVE) {
                    double hd;
                    if (dh <= 0.0) hd = x_pen > 0.0 ? 1.0 : 0.0;
                    else if (x_pen < -dh) hd = 0.0;
                    else if (x_pen > dh) hd = 1.0;
                    else hd = (x_pen + dh) / (2.0 * dh);
                    if (!(hd < 1e-30 && p_n < 1e-30)) {
                        double wm = hd * pp.k_pen;
                        double wg = dist > 1e-15 ? p_n / dist : 0.0;
#pragma unroll
                        for (int a = 0; a < 3; ++a)
#pragma unroll
                            for (int b = 0; b < 3; ++b) {
                                double nab = nv[a] * nv[b];
                                G[3 * a + b] = wm * nab - wg * ((a == b ? 1.0 : 0.0) - nab);
                            }
                    }
                }
                if ((modes & PAIR_TAN_F) && st != ST_INACTIVE && p_n > 0.0 && mu > 0.0) {
                    double q1 = P.z[i], q2 = P.z[cap + i];
                    double zn = sqrt(q1 * q1 + q2 * q2);
                    if (zn >= 1e-30) {
                        const double sgn = (modes & PAIR_TAN_C) ? -1.0 : 1.0;
                        double kt = P.k_t[i], qtn = P.qtn[i];
                        double D00, D01, D10, D11;
                        if (P.stick[i] || zn < 1e-30 || qtn < 1e-30) {
                            D00 = kt; D01 = 0.0; D10 = 0.0; D11 = kt;
                        } else {
                            double h1 = q1 / zn, h2 = q2 / zn;
                            double rk = ((mu * p_n) / qtn) * kt;
                            D00 = rk * (1.0 - h1 * h1); D01 = rk * (0.0 - h1 * h2);
                            D10 = rk * (0.0 - h2 * h1); D11 = rk * (1.0 - h2 * h2);
                        }
                        const double tv[2][3] = {{t1.x, t1.y, t1.z}, {t2.x, t2.y, t2.z}};
                        const double Dm[2][2] = {{D00, D01}, {D10, D11}};
                        double H[9];
#pragma unroll
                        for (int k = 0; k < 9; ++k) H[k] = 0.0;
#pragma unroll
                        for (int al = 0; al < 2; ++al)
#pragma unroll
                            for (int be = 0; be < 2; ++be) {
                                double dv = Dm[al][be];
                                if (fabs(dv) < 1e-30) continue;
#pragma unroll
                                for (int a = 0; a < 3; ++a)
#pragma unroll
                                    for (int b = 0; b < 3; ++b) H[3 * a + b] += dv * tv[al][a] * tv[be][b];
                            }
                        if (!(fabs(q1) < 1e-30 && fabs(q2) < 1e-30) && dist >= 1e-15) {
                            const double id = 1.0 / dist;
                            const double t2v[3] = {t2.x, t2.y, t2.z};
                            const double t1v[3] = {t1.x, t1.y, t1.z};
                            double M[9];
#pragma unroll
                            for (int a = 0; a < 3; ++a)
#pragma unroll
                                for (int b = 0; b < 3; ++b)
                                    M[3 * a + b] = -q1 * nv[a] * t1v[b] - q2 * t2v[a] * nv[b];
                            M[1] += q2 * t1.z;  M[2] += q2 * (-t1.y);
                            M[3] += q2 * (-t1.z); M[5] += q2 * t1.x;
                            M[6] += q2 * t1.y;  M[7] += q2 * (-t1.x);
#pragma unroll
                            for (int k = 0; k < 9; ++k) H[k] += id * M[k];
                        }
#pragma unroll
                        for (int k = 0; k < 9; ++k) G[k] += sgn * H[k];
                    }
                }
                if (coat) {
                    // coating stiffness k_eff n n^T (strategy.py:146-192), pairs with cc > 1e-15
                    if ((modes & PAIR_COAT_K) && cc > 1e-15) {
                        double k_eff = pp.coat_k;
                        if (pp.coat_c > 0.0 && pp.coat_dt > 0.0) k_eff += pp.coat_c / pp.coat_dt;
#pragma unroll
                        for (int a = 0; a < 3; ++a)
#pragma unroll
                            for (int b = 0; b < 3; ++b) G[3 * a + b] += k_eff * (nv[a] * nv[b]);
                    }
                    // coating friction stiffness T D T^T (strategy.py:288-370), pairs with cc > 0
                    if ((modes & PAIR_COAT_KF) && pp.coat_mu > 0.0 && cc > 0.0) {
                        const double kt = pp.coat_kt, pnc = pp.coat_k * cc;
                        const double q1 = P.cz[i], q2 = P.cz[cap + i];
                        const double zn = sqrt(q1 * q1 + q2 * q2), qtn = P.cqtn[i];
                        double D00, D01, D10, D11;
                        if (P.cstick[i] || zn < 1e-30 || qtn < 1e-30) {
                            D00 = kt; D01 = 0.0; D10 = 0.0; D11 = kt;
                        } else {
                            const double h1 = q1 / zn, h2 = q2 / zn;
                            const double rk = ((pp.coat_mu * pnc) / qtn) * kt;
                            D00 = rk * (1.0 - h1 * h1); D01 = rk * (0.0 - h1 * h2);
                            D10 = rk * (0.0 - h2 * h1); D11 = rk * (1.0 - h2 * h2);
                        }
                        const double T1[3] = {t1.x, t1.y, t1.z}, T2[3] = {t2.x, t2.y, t2.z};
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            // row a of T D = [T1[a] D00 + T2[a] D10, T1[a] D01 + T2[a] D11]
                            const double r0 = T1[a] * D00 + T2[a] * D10;
                            const double r1 = T1[a] * D01 + T2[a] * D11;
#pragma unroll
                            for (int b = 0; b < 3; ++b) G[3 * a + b] += r0 * T1[b] + r1 * T2[b];
                        }
                    }
                }
                GT* grec = sg + SG_REC * (size_t)si;
#pragma unroll
                for (int k = 0; k < 9; ++k) grec[k] = GT(G[k]);
                grec[SG_C + 0] = GT(1.0 - s);
                grec[SG_C + 1] = GT(s);
                grec[SG_C + 2] = GT(-(1.0 - t));
                grec[SG_C + 3] = GT(-t);
            }
            active_here = p_n > 0.0 ? 1 : 0;
        }
    }
done_pair:
    if (act_part) {
        // #(p_n > 0) of this workgroup: wave ballots, then one store per workgroup
        __shared__ int wc[BS / 64];
        unsigned long long m = __ballot(active_here);
        if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            int c = 0;
            for (int w = 0; w < BS / 64; ++w) c += wc[w];
            act_part[lb] = c;
        }
    }
}

// Replay, for pairs the fused assembly did not visit, of the geometry it
// would have given them: hist holds the translational displacement of each of
// the n_hist assemblies of a Newton call (3 values per node, assembly k at
// hist + k * stride).  Each pair of idx gets the narrowphase and the frame
// transport of every assembly in order, its state carried in registers, and
// the last one stored -- bitwise what one PAIR_GEOM pass per assembly leaves
// on a pair that contributes nothing (status INACTIVE, p_n = 0; no coating).
__global__ void __launch_bounds__(256)
pair_replay_kernel(int64_t n_idx, const int* __restrict__ idx, PairSoA P, const double* __restrict__ Xr,
                   const double* __restrict__ hist, int64_t stride, int n_hist) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_idx) return;
    const int64_t i = idx[w];
    const int64_t cap = P.cap;
    const int a0 = P.na0[i], a1 = P.na1[i], b0 = P.nb0[i], b1 = P.nb1[i];
    const double ra = P.ra[i], rb = P.rb[i];
    V3<double> pt1{P.t1[i], P.t1[cap + i], P.t1[2 * cap + i]};
    V3<double> pn{P.n[i], P.n[cap + i], P.n[2 * cap + i]};
    // reference positions stay in registers; each assembly's displacements are
    // loaded one step ahead, so their latency overlaps the previous geometry
    const int q[4] = {a0, a1, b0, b1};
    double X[12], cur[12], nxt[12];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) X[3 * k + c] = Xr[3 * q[k] + c];
    auto load = [&](int step, double* dst) {
        const double* h = hist + (size_t)step * stride;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) dst[3 * k + c] = h[3 * (size_t)q[k] + c];
    };
    if (n_hist > 0) load(0, cur);
    double s = 0.0, t = 0.0, gap = 0.0;
    for (int k = 0; k < n_hist; ++k) {
        if (k + 1 < n_hist) load(k + 1, nxt);
        V3<double> x[4];
#pragma unroll
        for (int m = 0; m < 4; ++m)
            x[m] = V3<double>{X[3 * m] + cur[3 * m], X[3 * m + 1] + cur[3 * m + 1], X[3 * m + 2] + cur[3 * m + 2]};
        double dist;
        V3<double> nrm, n, t1;
        closest_points(x[0], x[1], x[2], x[3], s, t, dist, nrm);
        gap = dist - (ra + rb);
        transport_frame(nrm, pn, pt1, n, t1);
        pn = n;
        pt1 = t1;
#pragma unroll
        for (int m = 0; m < 12; ++m) cur[m] = nxt[m];
    }
    if (n_hist <= 0) return;
    P.s[i] = s;
    P.t[i] = t;
    P.gap[i] = gap;
    P.n[i] = pn.x; P.n[cap + i] = pn.y; P.n[2 * cap + i] = pn.z;
    P.t1[i] = pt1.x; P.t1[cap + i] = pt1.y; P.t1[2 * cap + i] = pt1.z;
}

hipError_t launch_pair_replay(int64_t n_idx, const int* idx, const PairSoA& P, const double* Xr, const double* hist,
                              int64_t stride, int n_hist, hipStream_t st) {
    if (n_idx <= 0 || n_hist <= 0) return hipSuccess;
    hipLaunchKernelGGL(pair_replay_kernel, dim3((unsigned)((n_idx + 255) / 256)), dim3(256), 0, st, n_idx, idx, P,
                       Xr, hist, stride, n_hist);
    return hipGetLastError();
}

int pair_blocks(int64_t n_pairs, int bs) { return (int)((n_pairs + bs - 1) / bs); }

bool pair_block_ok(int bs) { return bs == 64 || bs == 128 || bs == 256 || bs == 512; }

template <class GT, int BS>
static void launch_pairs_bs(int64_t n_pairs, const PairSoA& P, const double* Xr, const double* u, const double* u_ref,
                            int modes, const PairParams& pp, double* nf, GT* sg, int* act_part, hipStream_t st,
                            const int* idx, bool g32) {
    dim3 blk(BS), grd((unsigned)xcd_padded(pair_blocks(n_pairs, BS)));
    if (pp.coat_on)
        hipLaunchKernelGGL((pair_kernel<GT, BS, true, false>), grd, blk, 0, st, n_pairs, P, Xr, u, u_ref, modes, pp,
                           nf, sg, act_part, idx);
    else if (g32)
        hipLaunchKernelGGL((pair_kernel<GT, BS, false, true>), grd, blk, 0, st, n_pairs, P, Xr, u, u_ref, modes, pp,
                           nf, sg, act_part, idx);
    else
        hipLaunchKernelGGL((pair_kernel<GT, BS, false, false>), grd, blk, 0, st, n_pairs, P, Xr, u, u_ref, modes, pp,
                           nf, sg, act_part, idx);
}

template <class GT>
static hipError_t launch_pairs_t(int bs, int64_t n_pairs, const PairSoA& P, const double* Xr, const double* u,
                                 const double* u_ref, int modes, const PairParams& pp, double* nf, GT* sg,
                                 int* act_part, hipStream_t st, const int* idx, bool g32) {
    switch (bs) {
        case 64: launch_pairs_bs<GT, 64>(n_pairs, P, Xr, u, u_ref, modes, pp, nf, sg, act_part, st, idx, g32); break;
        case 128: launch_pairs_bs<GT, 128>(n_pairs, P, Xr, u, u_ref, modes, pp, nf, sg, act_part, st, idx, g32); break;
        case 256: launch_pairs_bs<GT, 256>(n_pairs, P, Xr, u, u_ref, modes, pp, nf, sg, act_part, st, idx, g32); break;
        case 512: launch_pairs_bs<GT, 512>(n_pairs, P, Xr, u, u_ref, modes, pp, nf, sg, act_part, st, idx, g32); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The tangent records are FP64 in both precision forms: the device linear
// solve factors the pairs' blocks from them, and the FP32 rounding of contact
// blocks changes the Newton trajectory of frictional strands (measured at 91
// wires); the mixed form rounds them to FP32 only where the gather writes K.
// prec XK_PREC_FP32 (2) adds the FP32 geometry of non-contributing pairs.
hipError_t launch_pairs(int prec, int64_t n_pairs, const PairSoA& P, const double* Xr, const double* u,
                        const double* u_ref, int modes, const PairParams& pp, double* nf, void* sg,
                        int* act_part, hipStream_t st, const int* idx, int bs) {
    if (n_pairs <= 0) return hipSuccess;
    return launch_pairs_t<double>(bs, n_pairs, P, Xr, u, u_ref, modes, pp, nf, (double*)sg, act_part, st, idx,
                                  prec == 2);
}

}  // namespace xk
