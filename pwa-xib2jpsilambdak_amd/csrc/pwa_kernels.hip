// HIP kernels (gfx950) for the Lambda_b -> J/psi p K- amplitude likelihood.
//
//  k_prep   : raw helicity variables -> compact per-event planes (once per upload)
//  k_amp    : per event, the angular amplitudes: Lambda* + Zcs sum H0 and the
//             Pc angular-times-coupling sums S_r; stores the quadratic-form
//             coefficients (|H0|^2, p0_r, M_rr') and, for the gradient, H0, S_r
//  k_shape  : per event, the Pc lineshapes over the smearing sub-bins, |A|^2,
//             the per-event log-likelihood / normalisation term with a
//             compensated block reduction; for the gradient also the
//             lineshape / K-matrix parameter derivatives and BW1, BW2, a
//  k_adj    : per event, the coupling adjoint (Lambda*, Zcs, Pc couplings and
//             the Lambda* / Zcs lineshape parameters)
//  k_reduce, k_grad_reduce : deterministic compensated reductions of the
//             per-block partials
// The three per-event stages keep their intermediates in HBM (SoA planes),
// so each stage has a small register footprint and runs at several waves
// per SIMD instead of one register-bound fused kernel.
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

#include "pwa_adjoint.hpp"
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

// ---------------------------------------------------- intermediate planes
// pcs (quadratic form):  [H02, HNI, p0 (2 NZ), Md (NZ), Mo (NZ(NZ-1))]  -> 2 + 2 NZ + NZ^2
// amp (gradient only):   [H0 (16), S_r (16 NZ)]
// bws (gradient only):   [BW1 (2 NZ), B diag (NZ), B off (NZ(NZ-1)), a]  -> 1 + 2 NZ + NZ^2
__host__ __device__ constexpr int n_pcs(int nz) { return 2 + 2 * nz + nz * nz; }
__host__ __device__ constexpr int n_amp(int nz) { return 16 + 16 * nz; }
__host__ __device__ constexpr int n_bws(int nz) { return 1 + 2 * nz + nz * nz; }
__host__ __device__ constexpr int mo_index(int q, int r) { return r * (r - 1) / 2 + q; }  // q < r

// ------------------------------------------------------------ k_amp
template <int NZ, bool GRAD>
__global__ void __launch_bounds__(kBlock) k_amp(const double* __restrict__ ev, int64_t ld, int64_t n,
                                               const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                                               const C64* __restrict__ cpl, double* __restrict__ pcs,
                                               double* __restrict__ amp) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    auto P = [&](int p) -> double& { return pcs[(int64_t)p * ld + i]; };
    C64 H0[8];
    lambda_chain(S, D, cpl, ev, ld, i, ld_d(ev, ld, i, E_MKP), H0);
    double hni = 0.0;
#pragma unroll
    for (int h = 0; h < 8; ++h) hni += norm2(H0[h]);
    if (S->nX > 0) {
        C64 HX[8];
        zcs_chain(S, D, cpl, ev, ld, i, ld_d(ev, ld, i, E_MJPSIK), HX);
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            hni += norm2(HX[h]);
            H0[h] = H0[h] + HX[h];
        }
    }
    double h02 = 0.0;
#pragma unroll
    for (int h = 0; h < 8; ++h) h02 += norm2(H0[h]);
    P(0) = h02;
    P(1) = hni;
    if constexpr (GRAD) {
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            amp[(int64_t)(2 * h) * ld + i] = H0[h].re;
            amp[(int64_t)(2 * h + 1) * ld + i] = H0[h].im;
        }
    }
    if constexpr (NZ > 0) {
        const ChainAng a = load_chain(ev, ld, i, E_Z_CB);
        const double cP = ld_d(ev, ld, i, E_Z_CP), sP = ld_d(ev, ld, i, E_Z_SP);
        double cp[8], sp[8];
        powers(ld_d(ev, ld, i, E_Z_CZ), cp, 7);
        powers(ld_d(ev, ld, i, E_Z_SZ), sp, 7);
        C64 Sr[NZ][8];
#pragma unroll
        for (int r = 0; r < NZ; ++r) {
            pc_angular(S, cpl, S->nR + r, a, cp, sp, cP, sP, Sr[r]);
            C64 p0;
            double md = 0.0;
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                cmac(p0, conj(H0[h]), Sr[r][h]);
                md += norm2(Sr[r][h]);
            }
            P(2 + 2 * r) = p0.re;
            P(3 + 2 * r) = p0.im;
            P(2 + 2 * NZ + r) = md;
#pragma unroll
            for (int q = 0; q < r; ++q) {
                C64 m;
#pragma unroll
                for (int h = 0; h < 8; ++h) cmac(m, conj(Sr[q][h]), Sr[r][h]);
                const int o = 2 + 3 * NZ + 2 * mo_index(q, r);
                P(o) = m.re;
                P(o + 1) = m.im;
            }
            if constexpr (GRAD) {
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    amp[(int64_t)(16 + 16 * r + 2 * h) * ld + i] = Sr[r][h].re;
                    amp[(int64_t)(16 + 16 * r + 2 * h + 1) * ld + i] = Sr[r][h].im;
                }
            }
        }
    }
}

// ----------------------------------------------------------- k_shape
// Pc lineshapes over the event's sub-bins and the per-event likelihood term.
template <int NZ, bool MC, bool GRAD, bool KM>
__global__ void __launch_bounds__(kBlock) k_shape(const double* __restrict__ ev, int64_t ld, int64_t n, int scheme,
                                                 const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                                                 const double* __restrict__ pcs, double* __restrict__ bws,
                                                 double* __restrict__ fx_out, Partial* __restrict__ part, int noint,
                                                 GradLayout G, int NGp, double* __restrict__ gpart) {
    extern __shared__ double lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Emitter em{nullptr, nullptr, nullptr, 0, lane};
    if constexpr (GRAD) {
        em.stage = lds + wave * (kStage * 64);
        em.acc = lds + 4 * kStage * 64 + wave * NGp;
        em.sidx = reinterpret_cast<int32_t*>(lds + 4 * kStage * 64 + 4 * NGp) + wave * kStage;
        for (int q = lane; q < NGp; q += 64) em.acc[q] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    double s0 = 0, c0 = 0, s1 = 0, c1 = 0;
    const int nb = S->scheme_nbins[scheme];
    const double inb = 1.0 / nb;
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
        const int64_t gi = base + threadIdx.x;
        const bool valid = gi < n;
        const int64_t i = valid ? gi : n - 1;
        auto P = [&](int p) { return pcs[(int64_t)p * ld + i]; };
        const double w = ld_d(ev, ld, i, E_W);
        const double mjp = ld_d(ev, ld, i, E_MJPSIP);
        double f = P(0), fni = P(1);
        C64 p0[NZ > 0 ? NZ : 1], Mo[NZ > 0 ? NZ : 1][NZ > 0 ? NZ : 1];
        double Md[NZ > 0 ? NZ : 1];
        // lineshape-parameter sums (unweighted) for the gradient
        double gm[NZ > 0 ? NZ : 1], gw[NZ > 0 ? NZ : 1];
        C64 BW1[NZ > 0 ? NZ : 1], B2o[NZ > 0 ? NZ : 1][NZ > 0 ? NZ : 1];
        double B2d[NZ > 0 ? NZ : 1];
        double kgb[16], kfb[16], kmb[4];
        // the gradient is carried for at most one K-matrix class (the host
        // rejects models with more for pwa_eval_grad); the forward handles all
        const int ikc = (GRAD && S->km_nclass > 0) ? S->km_class[0] : -1;
        if constexpr (NZ > 0) {
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                p0[r] = C64(P(2 + 2 * r), P(3 + 2 * r));
                Md[r] = P(2 + 2 * NZ + r);
                gm[r] = 0.0;
                gw[r] = 0.0;
                BW1[r] = C64(0.0, 0.0);
                B2d[r] = 0.0;
#pragma unroll
                for (int q = 0; q < r; ++q) {
                    const int o = 2 + 3 * NZ + 2 * mo_index(q, r);
                    Mo[q][r] = C64(P(o), P(o + 1));
                    B2o[q][r] = C64(0.0, 0.0);
                }
            }
            if constexpr (GRAD) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    kgb[q] = 0.0;
                    kfb[q] = 0.0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) kmb[q] = 0.0;
            }
            const double sig = S->mres[0] + mjp * (S->mres[1] + mjp * S->mres[2]);
            double acc = 0.0;
            C64 L[NZ];
            for (int b = 0; b < nb; ++b) {
                const double mb = mjp + S->scheme_z[scheme][b] * sig;
                if constexpr (!GRAD) {
                    pc_lineshapes<NZ, KM>(S, D, mb, L);
                    acc += pc_form<NZ>(L, p0, Md, Mo);
                } else {
                    C64 Mt[NZ], dL0[NZ], dL1[NZ];
                    double Vx[NZ];
                    KMState st;
#pragma unroll
                    for (int r = 0; r < NZ; ++r) Mt[r] = C64(1.0, 0.0);
                    if constexpr (KM)
                    for (int c = 0; c < S->km_nclass; ++c) {
                        const int ik = S->km_class[c];
                        C64 Mtr[8];
                        if (ik == ikc) kmatrix_state(S, D, ik, mb, st, Mtr);
                        else kmatrix(S, D, ik, mb, Mtr);
#pragma unroll
                        for (int r = 0; r < NZ; ++r)
                            if (S->ik[S->nR + r] == ik) Mt[r] = select8(Mtr, S->ikm[S->nR + r]);
                    }
#pragma unroll
                    for (int r = 0; r < NZ; ++r) {
                        const int rr = S->nR + r;
                        const D2 m0 = D2::var(D->m0[rr], 0), g0 = D2::var(D->g0[rr], 1);
                        Cx<D2> Ls(D2(Mt[r].re), D2(Mt[r].im));
                        if (S->type[rr] == PWA_BREITWIGNER)
                            Ls = bw_twobody<D2>(S->P.m_p, S->P.m_jpsi, S->P.ffr, m0, g0, mb, S->L[rr]);
                        const D2 Vv = vertex_pc<D2>(S->P, S->spin2[rr], mb, m0);
                        Vx[r] = Vv.v;
                        Ls = Ls * Vv;
                        L[r] = cval(Ls);
                        dL0[r] = cder(Ls, 0);
                        dL1[r] = cder(Ls, 1);
                    }
                    acc += pc_form<NZ>(L, p0, Md, Mo);
                    C64 Ab[8];
#pragma unroll
                    for (int r = 0; r < NZ; ++r) {
                        BW1[r] = BW1[r] + L[r];
                        B2d[r] += norm2(L[r]);
#pragma unroll
                        for (int q = 0; q < r; ++q) B2o[q][r] = B2o[q][r] + L[q] * conj(L[r]);
                        // P_r = p0_r + sum_q conj(L_q) M[q][r]
                        C64 Pr = p0[r];
                        rmac(Pr, Md[r], conj(L[r]));
#pragma unroll
                        for (int q = 0; q < NZ; ++q) {
                            if (q < r) Pr = Pr + conj(L[q]) * Mo[q][r];
                            else if (q > r) Pr = Pr + conj(L[q]) * conj(Mo[r][q]);
                        }
                        gm[r] += re_mul(Pr, dL0[r]);
                        gw[r] += re_mul(Pr, dL1[r]);
                        if (ikc >= 0 && S->ik[S->nR + r] == ikc) {
                            const C64 add = Pr * Vx[r];
#pragma unroll
                            for (int q = 0; q < 8; ++q)
                                if (q == S->ikm[S->nR + r]) Ab[q] = Ab[q] + add;
                        }
                    }
                    if constexpr (KM)
                        if (ikc >= 0) kmatrix_adjoint(D, ikc, st, Ab, kgb, kfb, kmb);
                }
            }
            f += acc * inb;
            if (MC && noint) {
                if (scheme != 0) pc_lineshapes<NZ, KM>(S, D, mjp, L);
                C64 zero[NZ];
                fni += pc_form<NZ>(L, zero, Md, Mo);
            }
        }
        double v0, v1, aw;
        if (MC) {
            const bool unw = (w <= -1000.0);
            v0 = unw ? f : f * w;
            v1 = (noint && !unw) ? fni * w : 0.0;
            aw = unw ? 1.0 : w;
            if (valid && fx_out) fx_out[gi] = v0;
        } else {
            if (f > 0) v0 = log(f) * w;
            else v0 = (S->flags & PWA_F_NEG_FX_PENALTY) ? -1e12 : 0.0;
            v1 = (f > 0) ? 0.0 : 1.0;
            aw = (f > 0) ? w / f : 0.0;
            if (valid && fx_out) fx_out[gi] = f;
        }
        if (!valid) {
            v0 = 0.0;
            v1 = 0.0;
            aw = 0.0;
        }
        kadd(s0, c0, v0);
        kadd(s1, c1, v1);
        if constexpr (GRAD) {
            auto B = [&](int p) -> double& { return bws[(int64_t)p * ld + i]; };
            if constexpr (NZ > 0) {
                // stores for the coupling adjoint (BW1, BW2 = sums over sub-bins / nb)
                if (valid) {
#pragma unroll
                    for (int r = 0; r < NZ; ++r) {
                        B(2 * r) = BW1[r].re * inb;
                        B(2 * r + 1) = BW1[r].im * inb;
                        B(2 * NZ + r) = B2d[r] * inb;
#pragma unroll
                        for (int q = 0; q < r; ++q) {
                            const int o = 3 * NZ + 2 * mo_index(q, r);
                            B(o) = B2o[q][r].re * inb;
                            B(o + 1) = B2o[q][r].im * inb;
                        }
                    }
                }
                const double sc = 2.0 * aw * inb;
#pragma unroll
                for (int r = 0; r < NZ; ++r) {
                    em.emit(G.res(S->nR + r), sc * gm[r]);
                    em.emit(G.res(S->nR + r) + 1, sc * gw[r]);
                }
                if (KM && ikc >= 0) {
                    const int o = G.km(ikc);
                    for (int q = 0; q < 16; ++q) em.emit(o + q, sc * kgb[q]);
                    for (int q = 0; q < 16; ++q) em.emit(o + 16 + q, sc * kfb[q]);
                    for (int q = 0; q < 4; ++q) em.emit(o + 32 + q, sc * kmb[q]);
                }
            }
            if (valid) B(n_bws(NZ) - 1) = aw;
        }
    }
    if constexpr (GRAD) {
        if (em.cnt > 0) em.flush();
    }
    block_reduce2(s0, c0, s1, c1);
    if (threadIdx.x == 0) part[blockIdx.x] = Partial{s0, c0, s1, c1};
    if constexpr (GRAD) {
        __syncthreads();
        for (int q = threadIdx.x; q < G.ng; q += kBlock) {
            double t = 0.0;
#pragma unroll
            for (int wv = 0; wv < 4; ++wv) t += lds[4 * kStage * 64 + wv * NGp + q];
            gpart[(int64_t)blockIdx.x * G.ng + q] = t;
        }
    }
}

// ------------------------------------------------------------- k_adj
// Coupling adjoint: Y_h = H0_h + sum_r BW1_r S_rh for the Lambda* / Zcs
// couplings and lineshapes, U_rh = BW1_r conj(H0_h) + sum_q BW2_rq conj(S_qh)
// for the Pc couplings; every contribution weighted by the event's a.
template <int NZ>
__global__ void __launch_bounds__(kBlock) k_adj(const double* __restrict__ ev, int64_t ld, int64_t n,
                                               const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                                               const C64* __restrict__ cpl, const double* __restrict__ amp,
                                               const double* __restrict__ bws, GradLayout G, int NGp,
                                               double* __restrict__ gpart) {
    extern __shared__ double lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Emitter em{lds + wave * (kStage * 64),
               reinterpret_cast<int32_t*>(lds + 4 * kStage * 64 + 4 * NGp) + wave * kStage,
               lds + 4 * kStage * 64 + wave * NGp, 0, lane};
    for (int q = lane; q < NGp; q += 64) em.acc[q] = 0.0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
        const int64_t gi = base + threadIdx.x;
        const bool valid = gi < n;
        const int64_t i = valid ? gi : n - 1;
        auto A = [&](int p) { return amp[(int64_t)p * ld + i]; };
        auto B = [&](int p) { return bws[(int64_t)p * ld + i]; };
        const double a2 = valid ? 2.0 * B(n_bws(NZ) - 1) : 0.0;
        C64 H0[8], Y[8];
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            H0[h] = C64(A(2 * h), A(2 * h + 1));
            Y[h] = H0[h];
        }
        if constexpr (NZ > 0) {
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                const C64 bw1(B(2 * r), B(2 * r + 1));
#pragma unroll
                for (int h = 0; h < 8; ++h) cmac(Y[h], bw1, C64(A(16 + 16 * r + 2 * h), A(16 + 16 * r + 2 * h + 1)));
            }
        }
        lambda_adjoint(S, D, cpl, ev, ld, i, ld_d(ev, ld, i, E_MKP), Y, a2, G, em);
        if (S->nX > 0) zcs_adjoint(S, D, cpl, ev, ld, i, ld_d(ev, ld, i, E_MJPSIK), Y, a2, G, em);
        if constexpr (NZ > 0) {
            const ChainAng za = load_chain(ev, ld, i, E_Z_CB);
            const double cP = ld_d(ev, ld, i, E_Z_CP), sP = ld_d(ev, ld, i, E_Z_SP);
            double cp[8], sp[8];
            powers(ld_d(ev, ld, i, E_Z_CZ), cp, 7);
            powers(ld_d(ev, ld, i, E_Z_SZ), sp, 7);
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                const C64 bw1(B(2 * r), B(2 * r + 1));
                C64 U[8];
#pragma unroll
                for (int h = 0; h < 8; ++h) U[h] = bw1 * conj(H0[h]);
#pragma unroll
                for (int q = 0; q < NZ; ++q) {
                    // BW2_rq = (1/nb) sum_b L_r conj(L_q)
                    C64 b2;
                    if (q == r) b2 = C64(B(2 * NZ + r), 0.0);
                    else if (q < r) {
                        const int o = 3 * NZ + 2 * mo_index(q, r);
                        b2 = conj(C64(B(o), B(o + 1)));
                    } else {
                        const int o = 3 * NZ + 2 * mo_index(r, q);
                        b2 = C64(B(o), B(o + 1));
                    }
#pragma unroll
                    for (int h = 0; h < 8; ++h)
                        cmac(U[h], b2, conj(C64(A(16 + 16 * q + 2 * h), A(16 + 16 * q + 2 * h + 1))));
                }
                pc_coupling_adjoint(S, S->nR + r, za, cp, sp, cP, sP, U, a2, G, em);
            }
        }
    }
    if (em.cnt > 0) em.flush();
    __syncthreads();
    for (int q = threadIdx.x; q < G.ng; q += kBlock) {
        double t = 0.0;
#pragma unroll
        for (int wv = 0; wv < 4; ++wv) t += lds[4 * kStage * 64 + wv * NGp + q];
        gpart[(int64_t)blockIdx.x * G.ng + q] = t;
    }
}

// ---------------------------------------------------------- launchers
size_t grad_lds_bytes(int NG) {
    const int NGp = (NG + 1) & ~1;
    return sizeof(double) * (4 * kStage * 64 + 4 * NGp) + sizeof(int32_t) * 4 * kStage;
}

int inter_planes(int nz) { return n_pcs(nz) + n_amp(nz) + n_bws(nz); }

template <int NZ>
static hipError_t run_stages_t(const StageArgs& a, hipStream_t st) {
    const unsigned eblocks = (unsigned)((a.n + kBlock - 1) / kBlock);
    double* pcs = a.inter;
    double* amp = a.inter + (int64_t)n_pcs(NZ) * a.ld;
    double* bws = amp + (int64_t)n_amp(NZ) * a.ld;
    GradLayout G{a.ncpl, a.nres, a.ng};
    const int NGp = (a.ng + 1) & ~1;
    const size_t lds = a.grad ? grad_lds_bytes(a.ng) : 0;
    hipError_t e;
    if (a.grad) hipLaunchKernelGGL((k_amp<NZ, true>), dim3(eblocks), dim3(kBlock), 0, st, a.ev, a.ld, a.n, a.S, a.D, a.cpl, pcs, amp);
    else hipLaunchKernelGGL((k_amp<NZ, false>), dim3(eblocks), dim3(kBlock), 0, st, a.ev, a.ld, a.n, a.S, a.D, a.cpl, pcs, amp);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const unsigned sblocks = (unsigned)a.shape_blocks;
#define PWA_SHAPE(MC, GR, KM)                                                                                      \
    do {                                                                                                           \
        if (GR) {                                                                                                  \
            e = hipFuncSetAttribute((const void*)k_shape<NZ, MC, GR, KM>,                                          \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                         \
            if (e != hipSuccess) return e;                                                                         \
        }                                                                                                          \
        hipLaunchKernelGGL((k_shape<NZ, MC, GR, KM>), dim3(sblocks), dim3(kBlock), lds, st, a.ev, a.ld, a.n,        \
                           a.scheme, a.S, a.D, pcs, bws, a.fx, a.part, a.noint, G, NGp, a.gpart_shape);            \
    } while (0)
    if (a.km) {
        if (a.mc && a.grad) PWA_SHAPE(true, true, true);
        else if (a.mc) PWA_SHAPE(true, false, true);
        else if (a.grad) PWA_SHAPE(false, true, true);
        else PWA_SHAPE(false, false, true);
    } else {
        if (a.mc && a.grad) PWA_SHAPE(true, true, false);
        else if (a.mc) PWA_SHAPE(true, false, false);
        else if (a.grad) PWA_SHAPE(false, true, false);
        else PWA_SHAPE(false, false, false);
    }
#undef PWA_SHAPE
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.grad) {
        e = hipFuncSetAttribute((const void*)k_adj<NZ>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_adj<NZ>), dim3((unsigned)a.adj_blocks), dim3(kBlock), lds, st, a.ev, a.ld, a.n, a.S,
                           a.D, a.cpl, amp, bws, G, NGp, a.gpart_adj);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t run_stages(int nz, const StageArgs& a, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    switch (nz) {
#define PWA_NZ(k) case k: return run_stages_t<k>(a, st);
    PWA_NZ(0) PWA_NZ(1) PWA_NZ(2) PWA_NZ(3) PWA_NZ(4) PWA_NZ(5) PWA_NZ(6) PWA_NZ(7) PWA_NZ(8)
#undef PWA_NZ
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

hipError_t launch_reduce(const Partial* part, int ngroups, const int32_t* beg, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce, dim3(ngroups), dim3(kBlock), 0, st, part, beg, out);
    return hipGetLastError();
}

// out[k * NG + q] = sum over the partial rows [beg[k], beg[k+1]) of gpart[row * NG + q];
// 8 row-slices per component reduced in parallel, then combined in order.
__global__ void __launch_bounds__(kBlock) k_grad_reduce(const double* __restrict__ gpart, const int32_t* __restrict__ beg,
                                                         int NG, double* __restrict__ out) {
    __shared__ double red[8][32];
    const int q = blockIdx.x * 32 + (threadIdx.x & 31);
    const int slice = threadIdx.x >> 5;
    const int k = blockIdx.y;
    const int lo = beg[k], hi = beg[k + 1];
    double s = 0.0, c = 0.0;
    if (q < NG)
        for (int row = lo + slice; row < hi; row += 8) kadd(s, c, gpart[(int64_t)row * NG + q]);
    red[slice][threadIdx.x & 31] = s + c;
    __syncthreads();
    if (slice == 0 && q < NG) {
        double t = 0.0, tc = 0.0;
        for (int z = 0; z < 8; ++z) kadd(t, tc, red[z][threadIdx.x]);
        out[k * NG + q] = t + tc;
    }
}

hipError_t launch_grad_reduce(const double* gpart, const int32_t* beg, int NG, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_grad_reduce, dim3((NG + 31) / 32, 2), dim3(kBlock), 0, st, gpart, beg, NG, out);
    return hipGetLastError();
}

} // namespace pwa
