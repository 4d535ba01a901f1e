// Factorised lineshapes.  For the usual parameter ranges every lineshape of
// the model is (lineshape x Lambda_b-vertex barrier term)
//
//   L(m) = C * G(m) * 1 / (a - i b),   a = m0^2 - m^2,  b = K * E(m)      (Breit-Wigner)
//   L(m) = C * G(m)                                                        (vertex only)
//
// with per-resonance constants C, K, m0^2 (functions of m0, Gamma0 only,
// computed once per evaluation on the host together with their m0 / Gamma0
// derivatives) and per-event factors shared by every resonance of a chain
// with the same orbital momenta (LR, LB):
//
//   G(m) = pR^LR / sqrt(B_LR(z)) * pB^LB / sqrt(B_LB(zB))
//   E(m) = pR^(2 LR + 1) / (m B_LR(z))
//
// pR = breakup momentum of the resonance, pB = that of the Lambda_b vertex,
// z = (ffr pR)^2, zB = (ffb pB)^2, B_L the Blatt-Weisskopf polynomials.  This
// is the product of bw_lambda / bw_twobody and vertex_* (pwa_math.hpp) with the
// m0-dependent factors collected into C and K; the resonances whose branch
// conditions leave the usual ranges (below-threshold masses, the Sigma pi
// width, negative masses) keep the direct evaluation (kind LS_GENERIC): for
// the Lambda* and Zcs chains in a separate per-event pass (k_lsgen) whose
// values and derivatives the amplitude kernels read back, for the Pc chain
// inline.
#pragma once
#include <hip/hip_runtime.h>

#include "pwa_amp.h"
#include "pwa_math.hpp"

namespace pwa {

enum LsKind { LS_GENERIC = 0, LS_BW = 1, LS_VERTEX = 2 };

// Planes of the per-event generic-lineshape buffer: for slot k, planes
// 6k .. 6k+5 hold Re, Im of L and of dL/dm0, dL/dGamma0.
constexpr int kGenPlanes = 6;
enum Chain { CH_LAMBDA = 0, CH_PC = 1, CH_ZCS = 2 };

struct LsConst {
    int32_t kind, LR, LB;
    int32_t slot;  // LS_GENERIC Lambda* / Zcs: plane group of the per-event generic buffer
    double m2[3];  // m0^2 and its (m0, Gamma0) derivatives
    double K[3];
    double C[3];
};

// Orbital momentum of the Lambda_b vertex of each chain.
PWA_HD int vertex_LB(int chain, int spin2) {
    if (chain == CH_LAMBDA) return (spin2 <= 2) ? 0 : (spin2 - 3) / 2;
    if (chain == CH_PC) return (spin2 - 1) / 2;
    return spin2 >= 1 ? (spin2 - 1) / 2 : 0;
}

// Per-event kinematics of a chain at resonance mass m: s = m^2, 1/m, the
// breakup momenta pR (resonance -> daughters) and pB (Lambda_b -> resonance
// + bachelor), as breakup_d with the divisions by 2m and 2 m_lb folded into
// one reciprocal and a constant.
struct ChainKin {
    double s, imp, pR, pB;
};
PWA_HD ChainKin chain_kin(const Masses& P, int chain, double m) {
    double a1, a2, b;  // resonance daughters, Lambda_b bachelor
    if (chain == CH_LAMBDA) { a1 = P.m_p; a2 = P.m_k; b = P.m_jpsi; }
    else if (chain == CH_PC) { a1 = P.m_p; a2 = P.m_jpsi; b = P.m_k; }
    else { a1 = P.m_k; a2 = P.m_jpsi; b = P.m_p; }
    ChainKin k;
    k.s = m * m;
    k.imp = rcp_nr(m);
    const double u = (k.s - (a1 + a2) * (a1 + a2)) * (k.s - (a1 - a2) * (a1 - a2));
    k.pR = sqrt_nr(fabs(u)) * (0.5 * k.imp);
    const double mlb2 = P.m_lb * P.m_lb;
    const double v = (mlb2 - (b + m) * (b + m)) * (mlb2 - (b - m) * (b - m));
    k.pB = sqrt_nr(fabs(v)) * P.half_inv_mlb;
    return k;
}

PWA_HD bool has_barrier(int L) { return L >= 1 && L <= 5; }
PWA_HD double bpoly(double z, int L) { return has_barrier(L) ? bw_poly<double>(z, L) : 1.0; }

// Event factors for orbital momenta (LR, LB): G and E as above, and the
// vertex-only factor GB = pB^LB / sqrt(B_LB(zB)).
struct LsEv {
    int key;
    double G, E, GB;
};
PWA_HD LsEv ls_event(const Masses& P, const ChainKin& k, int LR, int LB, bool vertex_only) {
    const double pBL = ipow(k.pB, LB);
    double ibzB = 1.0;
    if (has_barrier(LB)) {
        const double rp = P.ffb * k.pB;
        ibzB = rcp_nr(bw_poly<double>(rp * rp, LB));
    }
    double G = 0.0, E = 0.0, GB = 0.0;
    if (vertex_only) {
        GB = has_barrier(LB) ? pBL * sqrt_nr(ibzB) : pBL;
    } else {
        double ibz = 1.0;
        if (has_barrier(LR)) {
            const double rr = P.ffr * k.pR;
            ibz = rcp_nr(bw_poly<double>(rr * rr, LR));
        }
        const double pL = ipow(k.pR, LR);
        G = (has_barrier(LR) || has_barrier(LB)) ? pL * pBL * sqrt_nr(ibz * ibzB) : pL * pBL;
        E = pL * pL * k.pR * ibz * k.imp;
    }
    LsEv e;
    e.key = vertex_only ? (64 + LB) : (LR * 8 + LB);
    e.G = G;
    e.E = E;
    e.GB = GB;
    return e;
}
// Refresh the cached event factors when the orbital momenta change.
PWA_HD void ls_event_cached(const Masses& P, const ChainKin& k, const LsConst& c, LsEv& e) {
    const bool vo = c.kind == LS_VERTEX;
    const int key = vo ? (64 + c.LB) : (c.LR * 8 + c.LB);
    if (key != e.key) e = ls_event(P, k, c.LR, c.LB, vo);
}

// Value of a fast lineshape.
PWA_HD C64 ls_value(const LsConst& c, const ChainKin& k, const LsEv& e) {
    if (c.kind == LS_VERTEX) return C64(c.C[0] * e.GB, 0.0);
    const double a = c.m2[0] - k.s;
    const double b = c.K[0] * e.E;
    const double t = (c.C[0] * e.G) * rcp_nr(a * a + b * b);
    return C64(a * t, b * t);
}

// Value and (m0, Gamma0) derivatives of a fast lineshape.
PWA_HD void ls_grad(const LsConst& c, const ChainKin& k, const LsEv& e, C64& v, C64& d0, C64& d1) {
    if (c.kind == LS_VERTEX) {
        v = C64(c.C[0] * e.GB, 0.0);
        d0 = C64(c.C[1] * e.GB, 0.0);
        d1 = C64(c.C[2] * e.GB, 0.0);
        return;
    }
    const double a = c.m2[0] - k.s;
    const double b = c.K[0] * e.E;
    const double inv = rcp_nr(a * a + b * b);
    const C64 w(a * inv, b * inv);  // 1 / (a - i b)
    const C64 w2 = w * w;
    const double cg = c.C[0] * e.G;
    const double t = cg * inv;  // the value exactly as ls_value forms it
    v = C64(a * t, b * t);
    // d w = w^2 (d a - i d b) * (-1)
    const C64 dw0 = w2 * C64(-c.m2[1], c.K[1] * e.E);
    const C64 dw1 = w2 * C64(-c.m2[2], c.K[2] * e.E);
    d0 = w * (c.C[1] * e.G) + dw0 * cg;
    d1 = w * (c.C[2] * e.G) + dw1 * cg;
}

// Two sub-bin masses at once: the key test and the kind test are uniform,
// so both evaluations sit in one branch and their dependency chains
// interleave.
PWA_HD void ls_event_cached2(const Masses& P, const ChainKin k[2], const LsConst& c, LsEv e[2]) {
    const bool vo = c.kind == LS_VERTEX;
    const int key = vo ? (64 + c.LB) : (c.LR * 8 + c.LB);
    if (key != e[0].key) {
        e[0] = ls_event(P, k[0], c.LR, c.LB, vo);
        e[1] = ls_event(P, k[1], c.LR, c.LB, vo);
    }
}
// Value (as ls_value forms it) and the parts of its (m0, Gamma0)
// derivatives: with w = 1 / (a - i b) (w = 1 for a vertex-only shape), the
// event factor g (G, or GB for a vertex-only shape) and E,
//   dL/dx = C_x g w + C g w^2 (-m2_x + i K_x E)       (x = m0, Gamma0),
// so that sum_b Re(P_b dL_b/dx) = C_x A1 + C (-m2_x A2 - K_x A3) with
//   A1 = sum_b g Re(P_b w),  A2 = sum_b g Re(P_b w^2),  A3 = sum_b g E Im(P_b w^2)
// (ls_grad_sums), which is what the smeared-event gradient needs of the
// derivatives.
struct LsParts {
    C64 w;
    double g, gE;
};
PWA_HD void ls_parts2(const LsConst& c, const ChainKin k[2], const LsEv e[2], C64 v[2], LsParts p[2]) {
    if (c.kind == LS_VERTEX) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            v[j] = C64(c.C[0] * e[j].GB, 0.0);
            p[j].w = C64(1.0, 0.0);
            p[j].g = e[j].GB;
            p[j].gE = 0.0;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double a = c.m2[0] - k[j].s;
        const double b = c.K[0] * e[j].E;
        const double inv = rcp_nr(a * a + b * b);
        const double t = (c.C[0] * e[j].G) * inv;
        v[j] = C64(a * t, b * t);
        p[j].w = C64(a * inv, b * inv);
        p[j].g = e[j].G;
        p[j].gE = e[j].G * e[j].E;
    }
}
// Accumulate the three sums above for one sub-bin with P-factor Pr.
PWA_HD void ls_grad_sums(const LsParts& p, C64 Pr, double& A1, double& A2, double& A3) {
    const C64 pw = Pr * p.w;
    const C64 pw2 = pw * p.w;
    A1 += p.g * pw.re;
    A2 += p.g * pw2.re;
    A3 += p.gE * pw2.im;
}
// (m0, Gamma0) derivative sums from A1..A3 (m2 and K are zero for a
// vertex-only shape, whose sums A2 = A1 and A3 = 0 then drop out).
PWA_HD void ls_grad_final(const LsConst& c, double A1, double A2, double A3, double& gm, double& gw) {
    gm = c.C[1] * A1 - c.C[0] * (c.m2[1] * A2 + c.K[1] * A3);
    gw = c.C[2] * A1 - c.C[0] * (c.m2[2] * A2 + c.K[2] * A3);
}

// Host: constants of resonance r for this evaluation (kind LS_GENERIC when
// the branch conditions of the direct evaluation leave the usual ranges).
// type: PWA_BREITWIGNER / PWA_NONRESONANT / PWA_KMATRIX (vertex term only).
inline LsConst ls_const(const Masses& P, int chain, int type, int spin2, int LR, double m0v, double g0v) {
    using D = Dual<2>;
    LsConst c{};
    c.slot = -1;
    c.LR = LR;
    c.LB = vertex_LB(chain, spin2);
    c.kind = LS_GENERIC;
    const int LB = c.LB;
    const D m0 = D::var(m0v, 0), g0 = D::var(g0v, 1);
    // vertex constant (normalisation at the resonance mass for Lambda*, at m_lb otherwise)
    double mB;
    double ma, mb;
    if (chain == CH_LAMBDA) { mB = P.m_jpsi; ma = P.m_p; mb = P.m_k; }
    else if (chain == CH_PC) { mB = P.m_k; ma = P.m_p; mb = P.m_jpsi; }
    else { mB = P.m_p; ma = P.m_k; mb = P.m_jpsi; }
    const bool vstd = (m0v > 1e-10 && m0v + mB < P.m_lb);
    if (LB > 0 && !vstd) return c;
    D cv(1.0);
    if (LB > 0) {
        const D pB0 = breakup_dm(P.m_lb, mB, m0);
        const D z0 = (P.ffb * pB0) * (P.ffb * pB0);
        const D b0 = (LB >= 1 && LB <= 5) ? bw_poly(z0, LB) : D(1.0);
        if (chain == CH_LAMBDA) cv = dsqrt(b0) / ipow(pB0, LB);
        else cv = dsqrt(b0) * (1.0 / ipow(P.m_lb, LB));
    }
    D C, K, m2;
    if (type == PWA_BREITWIGNER) {
        if (!(m0v > ma + mb)) return c;
        const D pR0 = breakup(m0, ma, mb);
        const D z0 = (P.ffr * pR0) * (P.ffr * pR0);
        const D b0 = (LR >= 1 && LR <= 5) ? bw_poly(z0, LR) : D(1.0);
        C = dsqrt(b0) / ipow(pR0, LR) * bw_normalisation(m0, g0) * cv;
        K = g0 * m0 * m0 * b0 / ipow(pR0, 2 * LR + 1);
        m2 = m0 * m0;
        c.kind = LS_BW;
    } else {
        C = cv;
        c.kind = LS_VERTEX;
    }
    c.C[0] = C.v; c.C[1] = C.d[0]; c.C[2] = C.d[1];
    c.K[0] = K.v; c.K[1] = K.d[0]; c.K[2] = K.d[1];
    c.m2[0] = m2.v; c.m2[1] = m2.d[0]; c.m2[2] = m2.d[1];
    return c;
}

} // namespace pwa
