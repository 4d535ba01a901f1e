// LS -> helicity coupling map of the Lambda_b -> J/psi p K- amplitude.
//
// Every helicity coupling c_j used by the kernels is a product of two affine
// forms in the complex fit parameters Z[a] = params[a] + i params[a+1]:
//     c_j = K_j * (sum_t alpha_jt Z[a_jt]) * B_j,
//     B_j = Z[b_j]                          (Lambda*: the resonance scale sc)
//     B_j = beta0_j + beta1_j Z[b_j]        (Pc: H_B; Zcs: H_Z)
// with real coefficients built once per model from Clebsch-Gordan values.
// This restates ConfigureCouplings<true>, PartWaveCoff, PartWaveCoffZ and
// PartWaveCoffX (src/RooDalitzAmplitude.cxx:2669-2986, 5532-5828), including
// the reference's rule that a coupling with |Z|^2 < 1e-10 is dropped.
// Because the map is explicit, its Jacobian is exact and cheap, which is how
// the analytic gradient is carried from helicity couplings to fit parameters.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

#include "pwa_amp.h"

namespace pwa {

using cplx = std::complex<double>;

// Fit-parameter layout: the RooArgList length of resonance r (a spline
// resonance also holds its knot triples, src/RooDalitzAmplitude.cxx:1721-1750)
// and the offset of its list in the concatenated vector.
inline int param_size(const pwa_model_desc& d, int r) {
    const bool spline = d.type[r] == PWA_MODELINDEPENDENT;
    if (r < d.n_lambda) return spline ? PWA_NPAR_LAMBDA_SPLINE : PWA_NPAR_LAMBDA;
    if (r < d.n_lambda + d.n_pc) return spline ? PWA_NPAR_PC_SPLINE : PWA_NPAR_PC;
    return PWA_NPAR_ZCS;
}
inline int param_offset(const pwa_model_desc& d, int r) {
    int off = 0;
    for (int q = 0; q < r; ++q) off += param_size(d, q);
    return off;
}

struct CouplingTerm {
    int a;          // index of Re Z in the fit-parameter vector
    double alpha;   // real coefficient
};

struct HelicityCoupling {
    double K = 1.0;
    std::vector<CouplingTerm> A;
    int b = -1;           // index of Re Z[b] (-1: B = beta0)
    bool b_scale = false; // B = Z[b] exactly (no constant, no skip)
    double beta0 = 0.0, beta1 = 0.0;
};

// <j1 m1; j2 m2 | J M>, arguments twice the quantum numbers (Racah formula).
double clebsch(int J, int M, int j1, int j2, int m1, int m2);

struct CouplingMap {
    std::vector<HelicityCoupling> c;   // reference _vamp order
    std::vector<int> ampcount;         // first helicity coupling of each resonance
    bool build(const pwa_model_desc& d, std::string& err);
    // value of every helicity coupling
    void eval(const double* params, cplx* out) const;
    // fit-parameter gradient from dNLL/dRe(c_j) + i dNLL/dIm(c_j) = gc[j]
    void chain(const double* params, const cplx* gc, double* grad) const;
};

} // namespace pwa
