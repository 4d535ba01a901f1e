// HIP kernels (gfx950) for the Lambda_b -> J/psi p K- amplitude likelihood.
//
//  k_prep   : raw helicity variables -> compact per-event planes (once per upload)
//  k_nll    : one thread per event; |A|^2 with the Pc lineshapes averaged
//             over the event's smearing sub-bins, fused with the per-event
//             log-likelihood / normalisation term and a compensated block
//             reduction (one Partial per block)
//  k_reduce : deterministic compensated reduction of the block partials
//
// Physics follows the reference's GetAmp2s / TemplateGetAmp2s
// (src/kernel_calEva.cu:85-811) and lineshape library
// (src/cu_DPFPropogator.cu), re-organised for the CDNA4 execution model:
//  * per-event angular tables are rebuilt in registers from 46 FP64 planes
//    (368 B/event) instead of reading ~2 KB/event of precomputed tables;
//  * Lambda* contributions are grouped by spin so each Wigner-d table is
//    built once per group; Pc/Zcs angular sums are factorised over the
//    (helicity, helicity) pairs that share a phase;
//  * the smeared Pc part is the quadratic form f_b = 2 Re(L.p0) + L^H M L
//    with p0, M formed once per event, so each sub-bin costs the lineshapes
//    plus O(R^2) flops;
//  * model parameters are read through uniform (scalar) loads.
#include <hip/hip_runtime.h>

#include "pwa_chains.hpp"

namespace pwa {

// -------------------------------------------------------------- prep
__global__ void __launch_bounds__(kBlock) k_prep(const double* __restrict__ raw, int64_t ldr, int64_t n,
                                                   double* __restrict__ ev, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    auto R = [&](int c) { return raw[c * ldr + i]; };
    auto W = [&](int p, double v) { ev[(int64_t)p * ld + i] = v; };
    auto half = [&](int pc, int ps, double x) {
        x = fmin(1.0, fmax(-1.0, x));
        W(pc, sqrt(0.5 * (1.0 + x)));
        W(ps, sqrt(0.5 * (1.0 - x)));
    };
    auto full = [&](int pc, int ps, double x) {
        x = fmin(1.0, fmax(-1.0, x));
        W(pc, x);
        W(ps, sqrt(fmax(0.0, (1.0 - x) * (1.0 + x))));
    };
    auto phase = [&](int pr, int pi, double phi) {
        double s, c;
        sincos(phi, &s, &c);
        W(pr, c);
        W(pi, s);
    };
    W(E_MKP, R(PWA_COL_MKP));
    W(E_MJPSIP, R(PWA_COL_MJPSIP));
    W(E_MJPSIK, R(PWA_COL_MJPSIK));
    W(E_W, R(PWA_NCOL));
    half(E_L_CB, E_L_SB, R(PWA_COL_COS_LB));
    full(E_L_CJ, E_L_SJ, R(PWA_COL_COS_JPSI));
    half(E_L_CL, E_L_SL, R(PWA_COL_COS_L));
    phase(E_L_EKR, E_L_EKI, 0.5 * R(PWA_COL_PHI_K));
    phase(E_L_EMR, E_L_EMI, R(PWA_COL_PHI_MU));
    half(E_Z_CB, E_Z_SB, R(PWA_COL_Z_COS_LB));
    full(E_Z_CJ, E_Z_SJ, R(PWA_COL_Z_COS_JPSI));
    half(E_Z_CZ, E_Z_SZ, R(PWA_COL_Z_COS_Z));
    half(E_Z_CP, E_Z_SP, R(PWA_COL_COS_P));
    phase(E_Z_EZR, E_Z_EZI, 0.5 * R(PWA_COL_Z_PHI_Z));
    phase(E_Z_EMR, E_Z_EMI, R(PWA_COL_Z_PHI_MU));
    phase(E_Z_EPR, E_Z_EPI, 0.5 * R(PWA_COL_Z_PHI_JPSI));
    phase(E_Z_EAR, E_Z_EAI, R(PWA_COL_ALPHA_MU));
    half(E_X_CB, E_X_SB, R(PWA_COL_X_COS_LB));
    full(E_X_CJ, E_X_SJ, R(PWA_COL_X_COS_JPSI));
    half(E_X_CZ, E_X_SZ, R(PWA_COL_X_COS_X));
    half(E_X_CP, E_X_SP, R(PWA_COL_X_COS_P));
    phase(E_X_EZR, E_X_EZI, 0.5 * R(PWA_COL_X_PHI_X));
    phase(E_X_EMR, E_X_EMI, R(PWA_COL_X_PHI_MU));
    phase(E_X_EPR, E_X_EPI, 0.5 * R(PWA_COL_X_PHI_JPSI));
    phase(E_X_EAR, E_X_EAI, R(PWA_COL_X_ALPHA_MU));
}

hipError_t launch_prep(const double* raw, int64_t ld_raw, int64_t n, double* ev, int64_t ld, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_prep, dim3((unsigned)blocks), dim3(kBlock), 0, st, raw, ld_raw, n, ev, ld);
    return hipGetLastError();
}

// ------------------------------------------------------------ k_nll
// One thread per event of a segment whose events share smearing scheme
// `scheme` (0 = unsmeared, a single sub-bin at z = 0).
template <int NZ, bool MC>
__global__ void __launch_bounds__(kBlock) k_nll(const double* __restrict__ ev, int64_t ld, int64_t n, int scheme,
                                               const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                                               const C64* __restrict__ cpl, double* __restrict__ fx_out,
                                               Partial* __restrict__ part, int noint) {
    const int64_t gi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = gi < n;
    const int64_t i = valid ? gi : n - 1;

    const double mkp = ld_d(ev, ld, i, E_MKP);
    const double mjp = ld_d(ev, ld, i, E_MJPSIP);
    const double w = ld_d(ev, ld, i, E_W);

    C64 HL[8];
    lambda_chain(S, D, cpl, ev, ld, i, mkp, HL);
    double hl2 = 0.0, hx2 = 0.0;
    C64 H0[8];
    if (S->nX > 0) {
        C64 HX[8];
        zcs_chain(S, D, cpl, ev, ld, i, ld_d(ev, ld, i, E_MJPSIK), HX);
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            hl2 += norm2(HL[h]);
            hx2 += norm2(HX[h]);
            H0[h] = HL[h] + HX[h];
        }
    } else {
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            hl2 += norm2(HL[h]);
            H0[h] = HL[h];
        }
    }
    double h02 = 0.0;
#pragma unroll
    for (int h = 0; h < 8; ++h) h02 += norm2(H0[h]);

    double f = h02, fni = hl2 + hx2;
    if constexpr (NZ > 0) {
        const ChainAng a = load_chain(ev, ld, i, E_Z_CB);
        const double cZ = ld_d(ev, ld, i, E_Z_CZ), sZ = ld_d(ev, ld, i, E_Z_SZ);
        const double cP = ld_d(ev, ld, i, E_Z_CP), sP = ld_d(ev, ld, i, E_Z_SP);
        double cp[8], sp[8];
        powers(cZ, cp, 7);
        powers(sZ, sp, 7);
        C64 Sr[NZ][8];
#pragma unroll
        for (int r = 0; r < NZ; ++r) pc_angular(S, cpl, S->nR + r, a, cp, sp, cP, sP, Sr[r]);
        C64 p0[NZ], Mo[NZ][NZ];
        double Md[NZ];
#pragma unroll
        for (int r = 0; r < NZ; ++r) {
            C64 acc;
            double d = 0.0;
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                cmac(acc, conj(H0[h]), Sr[r][h]);
                d += norm2(Sr[r][h]);
            }
            p0[r] = acc;
            Md[r] = d;
#pragma unroll
            for (int q = 0; q < r; ++q) {
                C64 m;
#pragma unroll
                for (int h = 0; h < 8; ++h) cmac(m, conj(Sr[q][h]), Sr[r][h]);
                Mo[q][r] = m;
            }
        }
        const int nb = S->scheme_nbins[scheme];
        const double sig = S->mres[0] + mjp * (S->mres[1] + mjp * S->mres[2]);
        double acc = 0.0;
        C64 L[NZ];
        for (int b = 0; b < nb; ++b) {
            const double mb = mjp + S->scheme_z[scheme][b] * sig;
            pc_lineshapes<NZ>(S, D, mb, L);
            acc += pc_form<NZ>(L, p0, Md, Mo);
        }
        f = h02 + acc * (1.0 / nb);
        if (MC && noint) {
            if (scheme != 0) pc_lineshapes<NZ>(S, D, mjp, L);
            C64 zero[NZ];
            fni += pc_form<NZ>(L, zero, Md, Mo);
        }
    }

    double v0 = 0.0, v1 = 0.0;
    if (MC) {
        const bool unw = (w <= -1000.0);
        const double fx = unw ? f : f * w;
        v0 = fx;
        v1 = (noint && !unw) ? fni * w : 0.0;
        if (valid && fx_out) fx_out[gi] = fx;
    } else {
        if (f > 0) v0 = log(f) * w;
        else v0 = (S->flags & PWA_F_NEG_FX_PENALTY) ? -1e12 : 0.0;
        v1 = (f > 0) ? 0.0 : 1.0;
        if (valid && fx_out) fx_out[gi] = f;
    }
    if (!valid) {
        v0 = 0.0;
        v1 = 0.0;
    }
    double c0 = 0.0, c1 = 0.0;
    block_reduce2(v0, c0, v1, c1);
    if (threadIdx.x == 0) part[blockIdx.x] = Partial{v0, c0, v1, c1};
}

template <int NZ>
static hipError_t launch_nll_t(bool mc, int scheme, const double* ev, int64_t ld, int64_t n, const DevStatic* S,
                               const DevDyn* D, const C64* cpl, double* fx, Partial* part, int noint, hipStream_t st) {
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    if (mc)
        hipLaunchKernelGGL((k_nll<NZ, true>), dim3(blocks), dim3(kBlock), 0, st, ev, ld, n, scheme, S, D, cpl, fx, part, noint);
    else
        hipLaunchKernelGGL((k_nll<NZ, false>), dim3(blocks), dim3(kBlock), 0, st, ev, ld, n, scheme, S, D, cpl, fx, part, noint);
    return hipGetLastError();
}

hipError_t launch_nll(int nZ, bool mc, int scheme, const double* ev, int64_t ld, int64_t n, const DevStatic* S,
                      const DevDyn* D, const C64* cpl, double* fx, Partial* part, int noint, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    switch (nZ) {
    case 0: return launch_nll_t<0>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 1: return launch_nll_t<1>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 2: return launch_nll_t<2>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 3: return launch_nll_t<3>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 4: return launch_nll_t<4>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 5: return launch_nll_t<5>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 6: return launch_nll_t<6>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 7: return launch_nll_t<7>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    case 8: return launch_nll_t<8>(mc, scheme, ev, ld, n, S, D, cpl, fx, part, noint, st);
    default: return hipErrorInvalidValue;
    }
}

// ----------------------------------------------------------- k_reduce
// Block b reduces the partials [beg[b], beg[b+1]) into out[4b .. 4b+3].
__global__ void __launch_bounds__(kBlock) k_reduce(const Partial* __restrict__ part, const int32_t* __restrict__ beg,
                                                    double* __restrict__ out) {
    const int b = blockIdx.x;
    const int lo = beg[b], hi = beg[b + 1];
    double s0 = 0, c0 = 0, s1 = 0, c1 = 0;
    for (int q = lo + threadIdx.x; q < hi; q += kBlock) {
        const Partial p = part[q];
        kadd(s0, c0, p.s0);
        c0 += p.c0;
        kadd(s1, c1, p.s1);
        c1 += p.c1;
    }
    block_reduce2(s0, c0, s1, c1);
    if (threadIdx.x == 0) {
        out[4 * b + 0] = s0;
        out[4 * b + 1] = c0;
        out[4 * b + 2] = s1;
        out[4 * b + 3] = c1;
    }
}

hipError_t launch_reduce(const Partial* part, int ngroups, const int32_t* beg, const int32_t* /*unused*/, double* out,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_reduce, dim3(ngroups), dim3(kBlock), 0, st, part, beg, out);
    return hipGetLastError();
}

} // namespace pwa
