// Host-side check of the dual-number derivatives of the triangle-cusp and
// revised K-matrix lineshapes (the same templates the kernels use): mass and
// width derivatives against central differences of the value, at points
// where the two step sizes agree (the cusp's complex-sqrt branch rule makes
// it discontinuous; a step across such a jump is skipped, not compared).
#include <cmath>
#include <cstdio>

#include "pwa_math.hpp"
using namespace pwa;
using D2 = Dual<2>;

int main() {
    const Masses P{5.61951, 0.938272046, 0.493677, 3.096916, 3.0, 3.0, 0.5 / 5.61951};
    double worst = 0.0;
    int n_cmp = 0, n_skip = 0;
    auto cmp = [&](const char* what, double an, double f1, double f2) {  // f1, f2: steps 2h, h
        if (fabs(f1 - f2) > 1e-5 * fmax(fabs(f1), 1e-3)) { ++n_skip; return; }  // a jump inside the step
        // Richardson: (4 f(h) - f(2h)) / 3 removes the O(h^2) truncation
        const double fd = (4.0 * f2 - f1) / 3.0;
        f2 = fd;
        const double e = fabs(an - f2) / fmax(fabs(f2), 1e-3);
        if (e > worst) worst = e;
        ++n_cmp;
        if (e > 1e-6) printf("%s an=% .10e fd=% .10e\n", what, an, f2);
    };
    // triangle cusp
    for (double m0 : {4.2, 4.3, 4.46}) {
        for (double g0 : {0.01, 0.05, 0.2}) {
            for (double mpp = 4.04; mpp < 5.12; mpp += 0.0537) {
                const Cx<D2> d = cusp_amp<D2>(P, D2::var(m0, 0), D2::var(g0, 1), mpp);
                for (int k = 0; k < 2; ++k) {
                    const double x0 = k == 0 ? m0 : g0;
                    auto f = [&](double h) {
                        const double a = k == 0 ? m0 + h : m0, b = k == 0 ? g0 : g0 + h;
                        const double c = k == 0 ? m0 - h : m0, e = k == 0 ? g0 : g0 - h;
                        const C64 p = cusp_amp<double>(P, a, b, mpp), q = cusp_amp<double>(P, c, e, mpp);
                        return C64((p.re - q.re) / (2 * h), (p.im - q.im) / (2 * h));
                    };
                    // step 1e-4 relative: the value carries ~1e-12 relative rounding
                    // noise (its cos-based square root), too much for smaller steps
                    const double h = 1e-4 * x0;
                    const C64 f1 = f(2 * h), f2 = f(h);
                    cmp("cusp re", d.re.d[k], f1.re, f2.re);
                    cmp("cusp im", d.im.d[k], f1.im, f2.im);
                }
            }
        }
    }
    // revised K-matrix: a 1/2- class with Lambda(1405) below threshold, two
    // poles and a non-resonant member
    const double mv[kRkmMem] = {1.405, 1.67, 1.80, -1.0}, gv[kRkmMem] = {0.05, 0.03, 0.2, 0.4};
    const uint32_t Lpk = 0;
    for (double mpp = 1.44; mpp < 2.52; mpp += 0.0711) {
        for (int r = 0; r < 4; ++r) {
            for (int j = 0; j < 4; ++j) {
                D2 m[kRkmMem], g[kRkmMem];
                for (int q = 0; q < kRkmMem; ++q) {
                    m[q] = q == j ? D2::var(mv[q], 0) : D2(mv[q]);
                    g[q] = q == j ? D2::var(gv[q], 1) : D2(gv[q]);
                }
                const Cx<D2> d = rkm_lambda<D2>(P, m, g, Lpk, 4, r, mpp);
                for (int k = 0; k < 2; ++k) {
                    auto f = [&](double h) {
                        double mp[kRkmMem], gp[kRkmMem], mm[kRkmMem], gm[kRkmMem];
                        for (int q = 0; q < kRkmMem; ++q) { mp[q] = mm[q] = mv[q]; gp[q] = gm[q] = gv[q]; }
                        (k == 0 ? mp : gp)[j] += h;
                        (k == 0 ? mm : gm)[j] -= h;
                        const C64 p = rkm_lambda<double>(P, mp, gp, Lpk, 4, r, mpp);
                        const C64 q = rkm_lambda<double>(P, mm, gm, Lpk, 4, r, mpp);
                        return C64((p.re - q.re) / (2 * h), (p.im - q.im) / (2 * h));
                    };
                    const double h = 1e-6 * fabs(k == 0 ? mv[j] : gv[j]);
                    const C64 f1 = f(2 * h), f2 = f(h);
                    cmp("rkm re", d.re.d[k], f1.re, f2.re);
                    cmp("rkm im", d.im.d[k], f1.im, f2.im);
                }
            }
        }
    }
    printf("compared %d, skipped %d (jump inside the step)\n", n_cmp, n_skip);
    printf("worst relative error %.3e\n", worst);
    return 0;
}
