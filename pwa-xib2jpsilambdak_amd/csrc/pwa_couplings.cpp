// LS -> helicity coupling map (see pwa_couplings.hpp).
#include "pwa_couplings.hpp"

#include <algorithm>
#include <string>

namespace pwa {

static double lfac(int n) { return std::lgamma(n + 1.0); }

double clebsch(int J, int M, int j1, int j2, int m1, int m2) {
    if (m1 + m2 != M) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
    if (J < std::abs(j1 - j2) || J > j1 + j2) return 0.0;
    if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1) || ((j1 + j2 + J) & 1)) return 0.0;
    // Racah: sqrt((2J+1) D(j1 j2 J) prod (j+-m)!) sum_k (-1)^k / (k! ...)
    const int a = (j1 + j2 - J) / 2, b = (j1 - j2 + J) / 2, c = (-j1 + j2 + J) / 2;
    const double lpre = 0.5 * (std::log(J + 1.0) + lfac(a) + lfac(b) + lfac(c) - lfac((j1 + j2 + J) / 2 + 1) +
                               lfac((j1 + m1) / 2) + lfac((j1 - m1) / 2) + lfac((j2 + m2) / 2) + lfac((j2 - m2) / 2) +
                               lfac((J + M) / 2) + lfac((J - M) / 2));
    const int kmin = std::max({0, (j2 - J - m1) / 2, (j1 + m2 - J) / 2});
    const int kmax = std::min({a, (j1 - m1) / 2, (j2 + m2) / 2});
    double s = 0.0;
    for (int k = kmin; k <= kmax; ++k) {
        const double l = lpre - lfac(k) - lfac(a - k) - lfac((j1 - m1) / 2 - k) - lfac((j2 + m2) / 2 - k) -
                         lfac((J - j2 + m1) / 2 + k) - lfac((J - j1 - m2) / 2 + k);
        s += ((k & 1) ? -1.0 : 1.0) * std::exp(l);
    }
    return s;
}

// <L/2 0; S/2 1/2 | 1/2 1/2> for 2S = 2*iS+1, 2L = 2S - 1 + 2*iL
// (the C2 table, src/RooDalitzAmplitude.cxx:96-107)
static double c2(int iS, int iL) {
    const int S = 2 * iS + 1;
    return clebsch(1, 1, S - 1 + 2 * iL, S, 0, 1);
}

static const int kHel1Two[6] = {1, 1, -1, -1, 3, -3};
static const int kHel2[6] = {1, 0, 0, -1, 1, -1};

// The adjoint kernels emit one row per coupling pair (k = 0, 1): they rely on
// the couplings of the second half of a resonance's list being those of the
// first half times one sign s (Lambda*: a parity sign; Pc, Zcs: s = 1).
static bool halves_match(const std::vector<HelicityCoupling>& c, size_t a0, bool lambda) {
    const size_t half = (c.size() - a0) / 2;
    if (2 * half != c.size() - a0 || half == 0) return false;
    const double s = c[a0].K / c[a0 + half].K;
    for (size_t j = 0; j < half; ++j) {
        const HelicityCoupling &h0 = c[a0 + j], &h1 = c[a0 + half + j];
        if (h0.K != s * h1.K || h0.b != h1.b || h0.b_scale != h1.b_scale || h0.beta0 != h1.beta0 ||
            h0.beta1 != h1.beta1 || h0.A.size() != h1.A.size())
            return false;
        for (size_t t = 0; t < h0.A.size(); ++t)
            if (h0.A[t].a != h1.A[t].a || h0.A[t].alpha != h1.A[t].alpha) return false;
    }
    return lambda ? (s == 1.0 || s == -1.0) : s == 1.0;
}

bool CouplingMap::build(const pwa_model_desc& d, std::string& err) {
    c.clear();
    ampcount.clear();
    const int n = d.n_lambda + d.n_pc + d.n_zcs;
    for (int r = 0; r < n; ++r) {
        const int off = param_offset(d, r);
        const int J = d.spin2[r];
        const int P = d.parity[r];
        ampcount.push_back((int)c.size());
        if (r < d.n_lambda) {
            // Lambda*: K = paritypart(k) * (-1)^(1-hel2); A = PartWaveCoff; B = sc
            if (J < 1 || J > 9 || (J % 2) != 1) { err = "Lambda* 2J must be odd in 1..9"; return false; }
            for (int k = 0; k <= 1; ++k) {
                const double paritypart = (k == 0) ? ((J == 1) ? -1.0 : (((J - 3) / 2) % 2 ? -1.0 : 1.0)) * P : 1.0;
                for (int m = 0; m < 6; ++m) {
                    if (J == 1 && m >= 4) continue;
                    const int l2 = kHel1Two[m], l3 = 2 * kHel2[m];
                    const int m1 = l2 - l3;
                    HelicityCoupling h;
                    h.K = paritypart * (((1 - kHel2[m]) % 2) ? -1.0 : 1.0);
                    int ib = 0;
                    for (int S = std::abs(J - 2); S <= J + 2; S += 2)
                        for (int L = S - 1; L <= S + 1; L += 2) {
                            double a = std::sqrt((L + 1.0) / 2.0) * clebsch(S, m1, J, 2, l2, -l3) * c2((S - 1) / 2, (L - S + 1) / 2);
                            if (m1 < 0 && L > S) a = -a;
                            h.A.push_back({off + 2 * ib, a});
                            ++ib;
                        }
                    // overall scale (scRe, scIm) at size - 7, after any spline knots
                    h.b = off + param_size(d, r) - 7;
                    h.b_scale = true;
                    c.push_back(h);
                }
            }
        } else if (r < d.n_lambda + d.n_pc) {
            // Pc: K = (-1)^(1-ihp); A = LS part of PartWaveCoffZ; B = H_B
            if (J < 1 || J > 7 || (J % 2) != 1) { err = "Pc 2J must be odd in 1..7"; return false; }
            for (int k = 0; k <= 1; ++k)
                for (int hpsi = -1; hpsi <= 1; ++hpsi)
                    for (int ihz = 0; ihz <= 1; ++ihz)
                        for (int ihp = 0; ihp <= 1; ++ihp) {
                            const int lz = 2 * ihz - 1, lp = 2 * ihp - 1, lpsi = 2 * hpsi;
                            if (std::abs(lpsi - lp) > J) continue;
                            HelicityCoupling h;
                            h.K = ((1 - ihp) % 2) ? -1.0 : 1.0;
                            const int m1 = lpsi - lp;
                            int ib = 1;
                            for (int S = 1; S <= 3; S += 2)
                                for (int L = std::abs(J - S); L <= J + S; L += 2) {
                                    if (((L / 2 + 1) % 2 ? -1 : 1) * P < 0) continue;
                                    if (S < std::abs(m1)) { ++ib; continue; }
                                    const double ccg1 = clebsch(S, m1, 2, 1, lpsi, -lp);
                                    double ccg;
                                    if (J == 1) {
                                        ccg = c2((S - 1) / 2, (L - S + 1) / 2);
                                        if (m1 < 0 && L > S) ccg = -ccg;
                                    } else {
                                        ccg = clebsch(J, m1, L, S, 0, m1);
                                    }
                                    h.A.push_back({off + 2 * ib, std::sqrt((L + 1.0) / (J + 1.0)) * ccg * ccg1});
                                    ++ib;
                                }
                            if (ib > 4) { err = "Pc coupling index out of range"; return false; }
                            h.b = off + 0;
                            h.beta0 = c2((J - 1) / 2, 0) * std::sqrt(J / 2.0);
                            h.beta1 = c2((J - 1) / 2, 1) * std::sqrt((J + 2) / 2.0) * (lz < 0 ? -1.0 : 1.0);
                            c.push_back(h);
                        }
        } else {
            // Zcs: K = (-1)^(1-ihp) (-1)^(1-helpsi); A = LS part of PartWaveCoffX; B = H_Z
            if (J < 0 || J > 6 || (J % 2) != 0) { err = "Zcs 2J must be even in 0..6"; return false; }
            for (int k = 0; k <= 1; ++k)
                for (int hpsi = -1; hpsi <= 1; ++hpsi) {
                    if (2 * std::abs(hpsi) > J) continue;
                    for (int ihp = 0; ihp <= 1; ++ihp)
                        for (int ihz = 0; ihz <= 1; ++ihz) {
                            const int lp = 2 * ihp - 1, lz = lp + (ihz ? 1 : -1), lpsi = 2 * hpsi;
                            if (std::abs(lz) > J) continue;
                            HelicityCoupling h;
                            h.K = (((1 - ihp) % 2) ? -1.0 : 1.0) * (((1 - hpsi) % 2) ? -1.0 : 1.0);
                            bool first = true;
                            h.beta0 = 0.0;
                            h.beta1 = 0.0;
                            for (int LZ = J - 2; LZ <= J + 2; LZ += 2) {
                                if (J == 0 && LZ != 2) continue;
                                if (LZ < 0) continue;
                                if (P * ((LZ / 2) % 2 ? -1 : 1) < 0) continue;
                                const double v = std::sqrt((LZ + 1.0) / (J + 1.0)) * clebsch(J, lpsi, LZ, 2, 0, lpsi);
                                if (first) { h.beta0 += v; first = false; }
                                else h.beta1 += v;
                            }
                            h.b = off + 8;
                            const int m1 = lz - lp;
                            int ib = 0;
                            for (int S = J - 1; S <= J + 1; S += 2) {
                                if (S < 0) continue;
                                for (int L = S - 1; L <= S + 1; L += 2) {
                                    double ccg1 = c2((S - 1) / 2, (L - S + 1) / 2);
                                    if (m1 < 0 && L > S) ccg1 = -ccg1;
                                    const double a = std::sqrt((L + 1.0) / 2.0) * clebsch(S, m1, J, 1, lz, -lp) * ccg1;
                                    h.A.push_back({off + 2 * ib, a});
                                    ++ib;
                                }
                            }
                            c.push_back(h);
                        }
                }
        }
        if (!halves_match(c, (size_t)ampcount.back(), r < d.n_lambda)) {
            err = "internal: the couplings of the two proton helicities differ by more than a sign";
            return false;
        }
    }
    return true;
}

static inline cplx Zat(const double* p, int a) { return cplx(p[a], p[a + 1]); }
static inline bool tiny(const cplx& z) { return std::norm(z) < 1e-10; }

void CouplingMap::eval(const double* p, cplx* out) const {
    for (size_t j = 0; j < c.size(); ++j) {
        const HelicityCoupling& h = c[j];
        cplx A(0.0, 0.0);
        for (const CouplingTerm& t : h.A) {
            const cplx z = Zat(p, t.a);
            if (!tiny(z)) A += t.alpha * z;
        }
        cplx B;
        if (h.b_scale) {
            B = Zat(p, h.b);
        } else {
            B = h.beta0;
            const cplx z = Zat(p, h.b);
            if (h.beta1 != 0.0 && !tiny(z)) B += h.beta1 * z;
        }
        out[j] = h.K * A * B;
    }
}

void CouplingMap::chain(const double* p, const cplx* gc, double* grad) const {
    for (size_t j = 0; j < c.size(); ++j) {
        const HelicityCoupling& h = c[j];
        cplx A(0.0, 0.0);
        for (const CouplingTerm& t : h.A) {
            const cplx z = Zat(p, t.a);
            if (!tiny(z)) A += t.alpha * z;
        }
        cplx B, dB(0.0, 0.0);
        if (h.b_scale) {
            B = Zat(p, h.b);
            dB = 1.0;
        } else {
            B = h.beta0;
            const cplx z = Zat(p, h.b);
            if (h.beta1 != 0.0 && !tiny(z)) {
                B += h.beta1 * z;
                dB = h.beta1;
            }
        }
        // dNLL/dZ (as Re + i Im) = gc * conj(dc/dZ)
        for (const CouplingTerm& t : h.A) {
            const cplx z = Zat(p, t.a);
            if (tiny(z)) continue;
            const cplx g = gc[j] * std::conj(h.K * t.alpha * B);
            grad[t.a] += g.real();
            grad[t.a + 1] += g.imag();
        }
        if (h.b >= 0 && dB != cplx(0.0, 0.0)) {
            const cplx g = gc[j] * std::conj(h.K * A * dB);
            grad[h.b] += g.real();
            grad[h.b + 1] += g.imag();
        }
    }
}

} // namespace pwa
