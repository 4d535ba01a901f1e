// Adjoint (gradient) building blocks: the LDS emitter that reduces per-event
// gradient contributions over a wave, the layout of the kernel-level gradient
// vector, and the adjoints of the Lambda*, Zcs and Pc coupling contractions.
#pragma once
#include <hip/hip_runtime.h>

#include "pwa_chains.hpp"

namespace pwa {

using D2 = Dual<2>;

constexpr int kStage = 16;  // components per staging tile

struct Emitter {
    double* stage;   // [kStage][64]
    int32_t* sidx;   // [kStage]
    double* acc;     // [NG]
    int cnt;
    int lane;
    __device__ __forceinline__ void flush() {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int c = lane & (kStage - 1), q = lane / kStage;  // 4 lanes per component, 16 events each
        double s = 0.0;
        if (c < cnt) {
#pragma unroll
            for (int e = 0; e < 16; ++e) s += stage[c * 64 + q * 16 + e];
        }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (q == 0 && c < cnt) atomicAdd(&acc[sidx[c]], s);  // a tile may hold one index twice
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        cnt = 0;
    }
    __device__ __forceinline__ void emit(int idx, double v) {
        stage[cnt * 64 + lane] = v;
        if (lane == 0) sidx[cnt] = idx;
        if (++cnt == kStage) flush();
    }
};

// Layout of the kernel-level gradient vector.
struct GradLayout {
    int ncpl;  // complex couplings: [0, 2 ncpl)
    int nres;  // (m0, width) pairs: [2 ncpl, 2 ncpl + 2 nres)
    int ng;    // total: 2 ncpl + 2 nres + 36 * PWA_KM_NCLASS
    __device__ __forceinline__ int cpl(int j) const { return 2 * j; }
    __device__ __forceinline__ int res(int r) const { return 2 * ncpl + 2 * r; }
    // per K-matrix class: g[16], f[16], pole masses [4]
    __device__ __forceinline__ int km(int ik) const { return 2 * ncpl + 2 * nres + 36 * ik; }
};

__device__ __forceinline__ C64 cval(const Cx<D2>& x) { return C64(x.re.v, x.im.v); }
__device__ __forceinline__ C64 cder(const Cx<D2>& x, int k) { return C64(x.re.d[k], x.im.d[k]); }

// ---------------------------------------------------------- Lambda* adjoint
static __device__ void lambda_adjoint(const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                               const C64* __restrict__ cpl, const double* __restrict__ ev, int64_t ld, int64_t i,
                               double mkp, const C64 Y[8], double a2, const GradLayout& G, Emitter& em) {
    const double cB = ld_d(ev, ld, i, E_L_CB), sB = ld_d(ev, ld, i, E_L_SB);
    const double cJ = ld_d(ev, ld, i, E_L_CJ), sJ = ld_d(ev, ld, i, E_L_SJ);
    const double cL = ld_d(ev, ld, i, E_L_CL), sL = ld_d(ev, ld, i, E_L_SL);
    const C64 eK = ld_c(ev, ld, i, E_L_EKR, E_L_EKI);
    const C64 eM = ld_c(ev, ld, i, E_L_EMR, E_L_EMI);
    const C64 eK3 = eK * eK * eK;
    const C64 ph[6] = {eK * eM, eK, conj(eK), conj(eK) * conj(eM), eK3 * eM, conj(eK3) * conj(eM)};
    const int l2[6] = {1, 0, 0, -1, 1, -1};
    const int dl2[6] = {-1, 1, -1, 1, 1, -1};
    // Q[k][m] = sum_ij A_ijm conj(Y_ijk)
    C64 Q[2][6];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            C64 acc;
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
                    rmac(acc, dhalf(2 * jj - 1, dl2[m], cB, sB) * done(l2[m], 2 * ii - 1, cJ, sJ),
                         conj(Y[4 * ii + 2 * jj + k]));
            Q[k][m] = ph[m] * acc;
        }
    double cp[10], sp[10];
    powers(cL, cp, 9);
    powers(sL, sp, 9);
    for (int g = 0; g < 5; ++g) {
        const int beg = S->lam_start[g], end = S->lam_start[g + 1];
        if (beg == end) continue;
        double dl[2][6];
        switch (g) {
        case 0: dl_table<1>(cp, sp, dl); break;
        case 1: dl_table<3>(cp, sp, dl); break;
        case 2: dl_table<5>(cp, sp, dl); break;
        case 3: dl_table<7>(cp, sp, dl); break;
        default: dl_table<9>(cp, sp, dl); break;
        }
        const int nm = (g == 0) ? 4 : 6;
        for (int q = beg; q < end; ++q) {
            const int r = S->lam_idx[q];
            const D2 m0 = D2::var(D->m0[r], 0), g0 = D2::var(D->g0[r], 1);
            Cx<D2> Ls(D2(1.0), D2(0.0));
            if (S->type[r] == PWA_BREITWIGNER) Ls = bw_lambda<D2>(S->P, m0, g0, mkp, S->L[r]);
            Ls = Ls * vertex_lambda<D2>(S->P, S->spin2[r], mkp, m0);
            const C64 Lv = cval(Ls);
            const C64* c = cpl + S->ampcount[r];
            C64 R;
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    if (m >= nm) continue;
                    const C64 dq = Q[k][m] * dl[k][m];
                    const C64 z = Lv * dq;
                    const int j = S->ampcount[r] + k * nm + m;
                    em.emit(G.cpl(j), a2 * z.re);
                    em.emit(G.cpl(j) + 1, -a2 * z.im);
                    cmac(R, c[k * nm + m], dq);
                }
            em.emit(G.res(r), a2 * re_mul(cder(Ls, 0), R));
            em.emit(G.res(r) + 1, a2 * re_mul(cder(Ls, 1), R));
        }
    }
}

// T[k][ps][d] = q_{ps,d} sum_ij p_ij dd_{ij,ps,d} W[4i+2j+k]  (adjoint of chain_contract)
__device__ __forceinline__ void chain_adjoint(const ChainAng& a, const C64 W[8], C64 T[2][3][2]) {
#pragma unroll
    for (int ps = 0; ps < 3; ++ps)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const C64 qv = cpow_int(a.eM, ps - 1) * (d ? a.eP : conj(a.eP));
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                C64 acc;
#pragma unroll
                for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 2; ++jj) {
                        const C64 pij = (jj ? a.eR : conj(a.eR)) * (ii ? a.eA : conj(a.eA));
                        const double dd = done(ps - 1, 2 * ii - 1, a.cJ, a.sJ) * dhalf(2 * jj - 1, 2 * d - 1, a.cB, a.sB);
                        cmac(acc, pij * dd, W[4 * ii + 2 * jj + k]);
                    }
                T[k][ps][d] = qv * acc;
            }
        }
}

// ------------------------------------------------------------- Zcs adjoint
static __device__ void zcs_adjoint(const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                            const C64* __restrict__ cpl, const double* __restrict__ ev, int64_t ld, int64_t i,
                            double mjk, const C64 Y[8], double a2, const GradLayout& G, Emitter& em) {
    const ChainAng a = load_chain(ev, ld, i, E_X_CB);
    const double cX = ld_d(ev, ld, i, E_X_CZ), sX = ld_d(ev, ld, i, E_X_SZ);
    const double cP = ld_d(ev, ld, i, E_X_CP), sP = ld_d(ev, ld, i, E_X_SP);
    double cp[8], sp[8];
    powers(cX, cp, 7);
    powers(sX, sp, 7);
    C64 cY[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) cY[h] = conj(Y[h]);
    C64 T[2][3][2];
    chain_adjoint(a, cY, T);
    for (int q = 0; q < S->nX; ++q) {
        const int r = S->nR + S->nZ + q;
        const int J2 = S->spin2[r];
        double dx[3][3];
        switch (J2) {
        case 0: dx_table<0>(cp, sp, dx); break;
        case 2: dx_table<2>(cp, sp, dx); break;
        case 4: dx_table<4>(cp, sp, dx); break;
        default: dx_table<6>(cp, sp, dx); break;
        }
        const D2 m0 = D2::var(D->m0[r], 0), g0 = D2::var(D->g0[r], 1);
        Cx<D2> Ls = bw_twobody<D2>(S->P.m_k, S->P.m_jpsi, S->P.ffr, m0, g0, mjk, S->L[r]);
        Ls = Ls * vertex_zcs<D2>(S->P, J2, mjk, m0);
        const C64 Lv = cval(Ls);
        const C64* c = cpl + S->ampcount[r];
        C64 R;
        int n = 0;
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int ps = 0; ps < 3; ++ps) {
                if (2 * abs(ps - 1) > J2) continue;
#pragma unroll
                for (int ip = 0; ip < 2; ++ip)
#pragma unroll
                    for (int iz = 0; iz < 2; ++iz) {
                        const int hz2 = (2 * ip - 1) + (2 * iz - 1);
                        if (hz2 > J2 || hz2 < -J2) continue;
                        const double w = dhalf(2 * ip - 1, 2 * k - 1, cP, sP) * dx[hz2 / 2 + 1][(1 - ps) + 1];
                        const C64 t = T[k][ps][iz] * w;
                        const C64 z = Lv * t;
                        const int j = S->ampcount[r] + n;
                        em.emit(G.cpl(j), a2 * z.re);
                        em.emit(G.cpl(j) + 1, -a2 * z.im);
                        cmac(R, c[n], t);
                        ++n;
                    }
            }
        em.emit(G.res(r), a2 * re_mul(cder(Ls, 0), R));
        em.emit(G.res(r) + 1, a2 * re_mul(cder(Ls, 1), R));
    }
}

// ---------------------------------------------------- Pc coupling adjoint
static __device__ void pc_coupling_adjoint(const DevStatic* __restrict__ S, int r, const ChainAng& a, const double* cp,
                                    const double* sp, double cP, double sP, const C64 U[8], double a2,
                                    const GradLayout& G, Emitter& em) {
    const int J2 = S->spin2[r];
    double dz[2][4];
    switch (J2) {
    case 1: dz_table<1>(cp, sp, dz); break;
    case 3: dz_table<3>(cp, sp, dz); break;
    case 5: dz_table<5>(cp, sp, dz); break;
    default: dz_table<7>(cp, sp, dz); break;
    }
    C64 T[2][3][2];
    chain_adjoint(a, U, T);
    int n = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int ps = 0; ps < 3; ++ps)
#pragma unroll
            for (int iz = 0; iz < 2; ++iz)
#pragma unroll
                for (int ip = 0; ip < 2; ++ip) {
                    const int n2 = 2 * (ps - 1) - (2 * ip - 1);
                    if (n2 > J2 || n2 < -J2) continue;
                    const double w = dhalf(2 * ip - 1, 2 * k - 1, cP, sP) * dz[iz][(n2 + 3) / 2];
                    const C64 z = T[k][ps][iz] * w;
                    const int j = S->ampcount[r] + n;
                    em.emit(G.cpl(j), a2 * z.re);
                    em.emit(G.cpl(j) + 1, -a2 * z.im);
                    ++n;
                }
}

} // namespace pwa
