// Gain formulas of the metric rules and the counter-based RNG, as plain
// scalar functions shared by the device kernels and the host C++ API.  The
// expression trees follow rules.cpp operator for operator, and everything
// is compiled with -ffp-contract=off, so each value is the IEEE-754 double
// the reference expression yields (int32 sub-products of the WCC gains wrap
// as they do there).
#ifndef PARCD_RULE_FORMULAS_H
#define PARCD_RULE_FORMULAS_H

#include <stdint.h>

#include "parcd.h"

#if defined(__HIPCC__)
#define PCD_HD __host__ __device__
#define PCD_ALIGN16 __align__(16)
#else
#define PCD_HD
#define PCD_ALIGN16 alignas(16)
#endif

namespace pcd {

// Community statistics as one 16-byte record: one gather per community.
struct PCD_ALIGN16 CommStats {
    int32_t size;
    int32_t int_vol;  // Partition::int_volume (2x internal edges)
    int32_t ext_vol;  // Partition::ext_volume
    int32_t pad;
};
static_assert(sizeof(CommStats) == 16, "community record is 16 bytes");

struct RuleParams {
    int kind;
    double lambda;
    int m;       // Graph::m
    double cc;   // Graph::clustering_coef (WCC)
    double inv_m;  // RN(1/m): device code's quotients by m read it (set_edges sets both)
};

PCD_HD inline void set_edges(RuleParams &r, int m) {
    r.m = m;
    r.inv_m = 1.0 / (double)m;
}

// x / m, rounded as the IEEE division rounds it, for an integer x and
// m = r.m (x and m below 2^31 here; the argument needs |x| < 2^53, m < 2^50).
// Device code replaces the division (a reciprocal iteration with scaling and
// fix-up steps, about ten instructions around a quarter-rate reciprocal) by
// one multiply and two fma, with y = RN(1/m):
//  * q0 = RN(x y) lies within 2 ulp of t = x/m;
//  * the residual x - q0 m is an integer multiple of ulp(q0) below 4m of
//    them, so the fma computes it exactly;
//  * q1 = RN(q0 + res y) = RN(t + eta) with |eta| < 2^-52 ulp(t).
// RN(t + eta) = RN(t): a rounding midpoint has 54 significant bits, which
// x/m cannot have (the odd 54-bit part would divide x < 2^53), and a point
// of the 2^-54 ulp(t) grid other than t lies at least 2^-54 ulp(t) / m away
// from t, farther than eta.  With m = 0 (no edges) every numerator is 0 and
// both forms give NaN.  Host code divides.
PCD_HD inline double div_m(const RuleParams &r, int x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double xd = (double)x, md = (double)r.m;
    const double q0 = xd * r.inv_m;
    const double res = __builtin_fma(-q0, md, xd);
    return __builtin_fma(res, r.inv_m, q0);
#else
    return x / (double)r.m;
#endif
}

PCD_HD inline int wrap_mul(int a, int b) { return (int)((unsigned)a * (unsigned)b); }

// rules.cpp:240-389 calculate_wcc_improvement
PCD_HD inline double wcc_improvement(int n_x, int n_y, int c_x, int c_y, int m_x, int m_y, int m_xy,
                                                  int b_x, int b_y, double cc) {
    double p_x = n_x > 1 ? m_x / (n_x * (n_x - 1.0) / 2.0) : 0.0;
    double p_y = n_y > 1 ? m_y / (n_y * (n_y - 1.0) / 2.0) : 0.0;
    double p_xy = c_x > 0 && c_y > 0 ? m_xy / (double)wrap_mul(c_x, c_y) : 0.0;
    double q_x = n_x > 0 ? (b_x - m_xy) / (double)n_x : 0;
    double q_y = n_y > 0 ? (b_y - m_xy) / (double)n_y : 0;
    double score = 0.0;
    const double t_x = (wrap_mul(n_x - 1, n_x - 2) / 2) * p_x * p_x * p_x;
    const double t_y = (wrap_mul(n_y - 1, n_y - 2) / 2) * p_y * p_y * p_y;
    if (c_x > 0) {
        double t_xy = t_x + wrap_mul(c_x - 1, c_y) * p_x * p_y * p_xy + (wrap_mul(c_y, c_y - 1) / 2) * p_xy * p_xy * p_y;
        double t_v = t_xy + q_x * (q_x - 1) / 2 * cc + q_x * (n_x - 1) * p_x * cc + c_y * q_x * p_xy * cc;
        double vt_xy = q_x;
        double vt_x = vt_xy + p_xy * c_y;
        double vt_v = vt_x + p_x * (n_x - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_x + n_y - 1 + vt_xy)) - (vt_v / t_v) * (t_x / (n_x - 1 + vt_x));
        if (t_v > 0) score += c_x * tr;
    }
    if (c_x < n_x) {
        double t_xy = t_x;
        double t_v = t_xy + q_x * (q_x - 1) / 2 * cc + q_x * (n_x - 1) * p_x * cc;
        double vt_xy = q_x;
        double vt_x = vt_xy;
        double vt_v = vt_x + p_x * (n_x - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_x + n_y - 1 + vt_xy)) - (vt_v / t_v) * (t_x / (n_x - 1 + vt_x));
        if (t_v > 0) score += (n_x - c_x) * tr;
    }
    if (c_y > 0) {
        double t_xy = t_y + wrap_mul(c_y - 1, c_x) * p_y * p_x * p_xy + (wrap_mul(c_x, c_x - 1) / 2) * p_xy * p_xy * p_x;
        double t_v = t_xy + q_y * (q_y - 1) / 2 * cc + q_y * (n_y - 1) * p_y * cc + c_x * q_y * p_xy * cc;
        double vt_xy = q_y;
        double vt_y = vt_xy + p_xy * c_x;
        double vt_v = vt_y + p_y * (n_y - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_y + n_x - 1 + vt_xy)) - (vt_v / t_v) * (t_y / (n_y - 1 + vt_y));
        if (t_v > 0) score += c_y * tr;
    }
    if (c_y < n_y) {
        double t_xy = t_y;
        double t_v = t_xy + q_y * (q_y - 1) / 2 * cc + q_y * (n_y - 1) * p_y * cc;
        double vt_xy = q_y;
        double vt_y = vt_xy;
        double vt_v = vt_y + p_y * (n_y - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_y + n_x - 1 + vt_xy)) - (vt_v / t_v) * (t_y / (n_y - 1 + vt_y));
        if (t_v > 0) score += (n_y - c_y) * tr;
    }
    return score;
}

// Rule kinds a caller is compiled for, one bit per kind.  Kernels built for
// modularity or CPM alone leave the WCC and conductance formulas (and the
// registers they take) out; RK_ALL dispatches on r.kind at run time.
constexpr unsigned RK_MOD = 1u << PCD_RULE_MODULARITY;
constexpr unsigned RK_CPM = 1u << PCD_RULE_CPM;
constexpr unsigned RK_WCC = 1u << PCD_RULE_WCC;
constexpr unsigned RK_COND = 1u << PCD_RULE_CONDUCTANCE;
constexpr unsigned RK_ALL = RK_MOD | RK_CPM | RK_WCC | RK_COND;

// Rule::score_join (rules.cpp:34-43, 136-139, 403-417, 93-102)
template <unsigned KM = RK_ALL>
PCD_HD inline double score_join(const RuleParams &r, int deg, CommStats c, int num) {
    if ((KM & RK_MOD) && (KM == RK_MOD || r.kind == PCD_RULE_MODULARITY)) {
        // x / (2m) is RN(x / m) / 2 exactly (no quotient here is subnormal)
        double e_ij = div_m(r, num);
        double a_i = div_m(r, c.int_vol + c.ext_vol) * 0.5;
        double a_j = div_m(r, deg) * 0.5;
        return e_ij - 2 * a_i * a_j;
    }
    if ((KM & RK_CPM) && (KM == RK_CPM || r.kind == PCD_RULE_CPM)) return num - r.lambda * c.size;
    if ((KM & RK_WCC) && (KM == RK_WCC || r.kind == PCD_RULE_WCC))
        return wcc_improvement(1, c.size, 1, num, 0, c.int_vol / 2, num, deg, c.ext_vol, r.cc);
    if (KM & RK_COND) {
        int s = c.size, i = c.int_vol, v = i + c.ext_vol;
        return wrap_mul(s + 1, i + 2 * num) / (double)(v + deg) - s * (i / (double)v);
    }
    return 0.0;  // kind outside KM: never called so
}

// Rule::score_leave (rules.cpp:45-56, 141-144, 419-436, 104-114)
template <unsigned KM = RK_ALL>
PCD_HD inline double score_leave(const RuleParams &r, int deg, CommStats c, int num) {
    if ((KM & RK_MOD) && (KM == RK_MOD || r.kind == PCD_RULE_MODULARITY)) {
        double e_ij = div_m(r, num);
        double a_i = div_m(r, c.int_vol + c.ext_vol - deg) * 0.5;
        double a_j = div_m(r, deg) * 0.5;
        return -(e_ij - 2 * a_i * a_j);
    }
    if ((KM & RK_CPM) && (KM == RK_CPM || r.kind == PCD_RULE_CPM)) return -(num - r.lambda * (c.size - 1));
    if ((KM & RK_WCC) && (KM == RK_WCC || r.kind == PCD_RULE_WCC))
        return -wcc_improvement(1, c.size - 1, 1, num, 0, c.int_vol / 2 - num, num, deg, c.ext_vol - deg + 2 * num,
                                r.cc);
    if (KM & RK_COND) {
        int s = c.size, i = c.int_vol, v = i + c.ext_vol;
        return wrap_mul(s - 1, i - 2 * num) / (double)(v - deg) - s * (i / (double)v);
    }
    return 0.0;
}

// Rule::score_merge (rules.cpp:24-32, 131-134, 391-401, 76-91); KM as for
// score_join: a build for one rule kind compiles that formula alone.
template <unsigned KM = RK_ALL>
PCD_HD inline double score_merge(const RuleParams &r, CommStats a, CommStats b, int links) {
    if ((KM & RK_MOD) && (KM == RK_MOD || r.kind == PCD_RULE_MODULARITY)) {
        double e_ij = div_m(r, links);
        double a_i = div_m(r, a.int_vol + a.ext_vol) * 0.5;
        double a_j = div_m(r, b.int_vol + b.ext_vol) * 0.5;
        return e_ij - 2 * a_i * a_j;
    }
    if ((KM & RK_CPM) && (KM == RK_CPM || r.kind == PCD_RULE_CPM)) return links - r.lambda * a.size * b.size;
    if ((KM & RK_WCC) && (KM == RK_WCC || r.kind == PCD_RULE_WCC)) {
        int cx = links < a.size ? links : a.size;
        int cy = links < b.size ? links : b.size;
        return wcc_improvement(a.size, b.size, cx, cy, a.int_vol / 2, b.int_vol / 2, links, a.ext_vol, b.ext_vol, r.cc);
    }
    if (KM & RK_COND) {
        int si = a.size, sj = b.size, ii = a.int_vol, ij = b.int_vol, vi = ii + a.ext_vol, vj = ij + b.ext_vol;
        return wrap_mul(si + sj, ii + ij + links) / (double)(vi + vj + links) - wrap_mul(si, ii) / (double)vi -
               wrap_mul(sj, ij) / (double)vj;
    }
    return 0.0;  // kind outside KM: never called so
}

// ------------------------------------------------------ counter-based RNG
PCD_HD inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
PCD_HD inline uint64_t hash3(uint64_t seed, uint32_t a, uint32_t b) {
    uint64_t x = (((uint64_t)a << 32) | (uint64_t)b) + 1ULL;
    return mix64(seed + 0x9E3779B97F4A7C15ULL * x);
}
// refinement.cpp:116,129: active iff rand <= (int)(min(frac,1) * RAND_MAX);
// rand is the top 31 bits of hash3(seed, step, v).
PCD_HD inline bool selected(uint64_t seed, int step, int v, int f) {
    return (int)(uint32_t)(hash3(seed, (uint32_t)step, (uint32_t)v) >> 33) <= f;
}
// Tie key: per-(step, vertex) base hash, then mixed with the label.
PCD_HD inline uint64_t tie_base(uint64_t seed, int step, int v) {
    return hash3(seed ^ 0x5851F42D4C957F2DULL, (uint32_t)step, (uint32_t)v);
}
PCD_HD inline uint32_t tie_key(int tie, uint64_t base, int label) {
    if (tie == PCD_TIE_LOWEST) return (uint32_t)label;
    return (uint32_t)(mix64(base ^ ((uint64_t)(uint32_t)label * 0xD6E8FEB86659FD93ULL)) >> 32);
}
PCD_HD inline uint64_t merge_seed(uint64_t seed, int round) {
    return mix64(seed ^ (0xA0761D6478BD642FULL + (uint64_t)round));
}

}  // namespace pcd

#endif  // PARCD_RULE_FORMULAS_H
