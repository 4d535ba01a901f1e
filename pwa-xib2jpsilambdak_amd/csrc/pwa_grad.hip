// Analytic gradient of the amplitude likelihood (gfx950).
//
// k_grad evaluates, per event, |A|^2 and the derivative of a * |A|^2 with
// respect to every kernel-level parameter (complex helicity couplings,
// resonance masses and widths, K-matrix couplings / background terms / pole
// masses), where a = sw/|A|^2 for data and a = w for MC.  The host combines
//   dNLL/dtheta = sumW_data * G_mc / I - G_data.
//
// Derivatives:
//  * f = sum_h |H0|^2 + 2 Re(sum_r BW1_r p0_r) + sum_rr' BW2_rr' M_r'r  (smeared form)
//    dH0 -> adjoint Y_h = H0_h + sum_r BW1_r S_rh; dS_r -> U_rh = BW1_r conj(H0_h)
//    + sum_r' BW2_rr' conj(S_r'h); dL_r(b) -> P_r(b) = p0_r + sum_r' conj(L_r'(b)) M_r'r.
//  * Breit-Wigner and barrier factors: forward-mode duals in (m0, width).
//  * K-matrix: reverse mode through T = row 0 of (D(1 - iK rho))^-1, reusing
//    the LU factors (adjoint solve with the transposed factors).
// Reduction: every lane emits its event's contributions one component at a
// time (uniform order across the wave) into an LDS staging tile of
// 16 components x 64 lanes; a full tile is column-summed and added to the
// wave's LDS accumulator.  Blocks grid-stride over the events and write one
// partial gradient each; k_grad_reduce sums the partials in a fixed order.
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
        if (q == 0 && c < cnt) acc[sidx[c]] += s;
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
    __device__ __forceinline__ int cpl(int j) const { return 2 * j; }
    __device__ __forceinline__ int res(int r) const { return 2 * ncpl + 2 * r; }
    // per K-matrix class: g[16], f[16], pole masses [4]
    __device__ __forceinline__ int km(int ik) const { return 2 * ncpl + 2 * nres + 36 * ik; }
};

__device__ __forceinline__ C64 cval(const Cx<D2>& x) { return C64(x.re.v, x.im.v); }
__device__ __forceinline__ C64 cder(const Cx<D2>& x, int k) { return C64(x.re.d[k], x.im.d[k]); }
__device__ __forceinline__ double re_mul(const C64& a, const C64& b) { return a.re * b.re - a.im * b.im; }

// ---------------------------------------------------------- Lambda* adjoint
__device__ void lambda_adjoint(const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
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
            for (int k = 0; k < 2; ++k)
                for (int m = 0; m < nm; ++m) {
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
__device__ void zcs_adjoint(const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
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
        for (int k = 0; k < 2; ++k)
            for (int ps = 0; ps < 3; ++ps) {
                if (2 * abs(ps - 1) > J2) continue;
                for (int ip = 0; ip < 2; ++ip)
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
__device__ void pc_coupling_adjoint(const DevStatic* __restrict__ S, int r, const ChainAng& a, const double* cp,
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
    for (int k = 0; k < 2; ++k)
        for (int ps = 0; ps < 3; ++ps)
            for (int iz = 0; iz < 2; ++iz)
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

// ------------------------------------------------------ K-matrix adjoint
// Forward state of one class at one mass, kept for the reverse pass.
struct KMState {
    C64 rho[4];
    double den[4];
    bool clamped[4];
    double ms[4];
    double Dall;
    double kma[4][4];
    C64 Lm[4][4], U[4][4], iU[4];
    C64 T[4];
    double na;
    int np;
    double mfit[4], rcim[4], drcim[4];  // pole mass, Im rho_shift, d Im rho_shift / d mfit
    int ch[4];
};

// rho for the complex-mass shift channel and d(Im rho)/dM2 (branch as in csqrt_branch)
__device__ __forceinline__ void rho_cmass_d(double M2, double m1, double m2, double hw, double& im, double& dim) {
    const C64 a(m1 + m2, -hw), b(m1 - m2, hw);
    const C64 c1 = a * a, c2 = b * b;
    const C64 P = (c1 - C64(M2, 0.0)) * (c2 - C64(M2, 0.0));
    const C64 r = csqrt_branch(P);
    im = r.im / M2;
    if (r.re == 0.0 && r.im == 0.0) {
        dim = 0.0;
        return;
    }
    const C64 dP = C64(2.0 * M2, 0.0) - c1 - c2;
    const C64 dr = dP * cinv(r * 2.0);  // d sqrt(P) / dM2
    dim = dr.im / M2 - r.im / (M2 * M2);
}

__device__ void kmatrix_state(const DevStatic* __restrict__ S, const DevDyn* __restrict__ D, int ik, double mpp,
                              KMState& st, C64 Mtr[8]) {
    const int Lna = (ik >= 2) ? 1 : 0;
    const double s = mpp * mpp;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        st.rho[c] = S->km_cplx[ik][c] ? rho_cmass(s, S->km_ma[ik][c], S->km_mb[ik][c], S->km_hw[ik][c])
                                      : rho_real(s, S->km_ma[ik][c], S->km_mb[ik][c], Lna, S->P.ffb);
    const double* g = D->km[ik];
    const double* f = D->km[ik] + 16;
    const int np = D->km_npole[ik];
    st.np = np;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        st.den[p] = 1.0;
        st.clamped[p] = true;
        st.ch[p] = 0;
        st.mfit[p] = 0.0;
        st.rcim[p] = 0.0;
        st.drcim[p] = 0.0;
        if (p >= np) continue;
        const int r = D->km_pole_res[ik][p];
        const double mfit = D->m0[r];
        const int ch = D->km_pole_shift[ik][p];
        st.mfit[p] = mfit;
        st.ch[p] = ch;
        double m02 = mfit * mfit;
        if (ch > 0) {
            double im, dim;
            rho_cmass_d(m02, S->km_shift_ma[ch], S->km_shift_mb[ch], S->km_shift_hw[ch], im, dim);
            const double gc = g[p * 4 + ch];
            const double m0 = sqrt(m02 - gc * gc * im);
            st.rcim[p] = im;
            st.drcim[p] = dim * 2.0 * mfit;
            m02 = m0 * m0;
        }
        double d = m02 - s;
        bool cl = false;
        if ((S->flags & PWA_F_KM_CLAMP) && fabs(d) < 1e-3) {
            d = 1e-3;
            cl = true;
        }
        st.den[p] = d;
        st.clamped[p] = cl;
    }
    double Dall = 1.0;
#pragma unroll
    for (int p = 0; p < 4; ++p) Dall *= (p < np) ? st.den[p] : 1.0;
    st.Dall = Dall;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        double t = 1.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q != p && q < np) t *= st.den[q];
        st.ms[p] = t;
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = x; y < 4; ++y) {
            double v = f[x * 4 + y] * Dall;
#pragma unroll
            for (int p = 0; p < 4; ++p)
                if (p < np) v += g[p * 4 + x] * (g[p * 4 + y] * st.ms[p]);
            st.kma[x][y] = v;
            st.kma[y][x] = v;
        }
    C64 A[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const C64 t = st.rho[x] * st.kma[y][x];
            A[x][y] = C64(t.im, -t.re);
            if (x == y) A[x][y].re += Dall;
        }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
#pragma unroll
        for (int col = x; col < 4; ++col) {
            C64 u = A[x][col];
#pragma unroll
            for (int k = 0; k < x; ++k) u = u - st.Lm[x][k] * st.U[k][col];
            st.U[x][col] = u;
        }
        st.iU[x] = cinv(st.U[x][x]);
#pragma unroll
        for (int row = x + 1; row < 4; ++row) {
            C64 l = A[row][x];
#pragma unroll
            for (int k = 0; k < x; ++k) l = l - st.Lm[row][k] * st.U[k][x];
            st.Lm[row][x] = l * st.iU[x];
        }
    }
    C64 y[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        C64 v = (x == 0) ? C64(1.0, 0.0) : C64(0.0, 0.0);
#pragma unroll
        for (int b = 0; b < x; ++b) v = v - st.Lm[x][b] * y[b];
        y[x] = v;
    }
#pragma unroll
    for (int x = 3; x >= 0; --x) {
        C64 v = y[x];
#pragma unroll
        for (int b = x + 1; b < 4; ++b) v = v - st.U[x][b] * st.T[b];
        st.T[x] = st.iU[x] * v;
    }
    st.na = Lna ? fbw2(s, S->km_ma[ik][0], S->km_mb[ik][0], Lna, S->P.ffb) : 1.0;
#pragma unroll
    for (int p = 0; p < 8; ++p) Mtr[p] = C64(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= np) continue;
        C64 acc;
#pragma unroll
        for (int c = 0; c < 4; ++c) rmac(acc, g[p * 4 + c] * st.ms[p], st.T[c]);
        Mtr[p] = acc * st.na;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const C64 v = st.T[c] * (Dall * st.na);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q == np + c) Mtr[q] = v;
    }
}

// Reverse pass: Abar[q] is the adjoint of Mtr[q] (delta F = Re sum_q Abar_q dMtr_q).
// Accumulates Re-parts into gb[16] (g[p][c]), fb[16] (f[x][y], x <= y), mb[4] (pole masses).
__device__ void kmatrix_adjoint(const DevDyn* __restrict__ D, int ik, const KMState& st, const C64 Abar[8],
                                double gb[16], double fb[16], double mb[4]) {
    const double* g = D->km[ik];
    const double* f = D->km[ik] + 16;
    const int np = st.np;
    C64 Tb[4];
    C64 Db(0.0, 0.0), msb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        C64 an;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q == np + c) an = Abar[q];
        Tb[c] = an * (st.na * st.Dall);
        cmac(Db, an * st.na, st.T[c]);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= np) continue;
        const C64 ap = Abar[p] * st.na;
        C64 gt;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            rmac(Tb[c], g[p * 4 + c] * st.ms[p], ap);
            gb[p * 4 + c] += re_mul(ap, st.T[c]) * st.ms[p];
            rmac(gt, g[p * 4 + c], st.T[c]);
        }
        msb[p] = ap * gt;
    }
    // lambda: A^T lambda = Tb with A = L U  ->  U^T y = Tb, L^T lambda = y
    C64 yv[4], lam[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        C64 v = Tb[x];
#pragma unroll
        for (int b = 0; b < x; ++b) v = v - st.U[b][x] * yv[b];
        yv[x] = v * st.iU[x];
    }
#pragma unroll
    for (int x = 3; x >= 0; --x) {
        C64 v = yv[x];
#pragma unroll
        for (int b = x + 1; b < 4; ++b) v = v - st.Lm[b][x] * lam[b];
        lam[x] = v;
    }
    // Abar_matrix[x][y] = -lam_x T_y; A[x][y] = Dall delta - i kma[y][x] rho[x]
    // kma adjoint (row y, col x): kbar[y][x] = -i rho[x] * Abar_matrix[x][y] = i rho[x] lam_x T_y
#pragma unroll
    for (int x = 0; x < 4; ++x) Db = Db - lam[x] * st.T[x];
    C64 ub[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const C64 t = st.rho[y] * lam[y] * st.T[x];  // kbar[x][y] = i rho[y] lam_y T_x
            ub[x][y] = C64(-t.im, t.re);
        }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = x; y < 4; ++y) {
            const C64 u = (x == y) ? ub[x][x] : ub[x][y] + ub[y][x];
            fb[x * 4 + y] += u.re * st.Dall;
            rmac(Db, f[x * 4 + y], u);
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                if (p >= np) continue;
                gb[p * 4 + x] += u.re * g[p * 4 + y] * st.ms[p];
                gb[p * 4 + y] += u.re * g[p * 4 + x] * st.ms[p];
                rmac(msb[p], g[p * 4 + x] * g[p * 4 + y], u);
            }
        }
    // D = prod den, ms_p = prod_{q != p} den
    double denb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q >= np) continue;
        double v = Db.re * st.ms[q];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (p >= np || p == q) continue;
            double t = 1.0;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (r < np && r != p && r != q) t *= st.den[r];
            v += msb[p].re * t;
        }
        denb[q] = st.clamped[q] ? 0.0 : v;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (p >= np) continue;
        const int ch = st.ch[p];
        if (ch > 0) {
            const double gc = g[p * 4 + ch];
            mb[p] += denb[p] * (2.0 * st.mfit[p] - gc * gc * st.drcim[p]);
            gb[p * 4 + ch] += denb[p] * (-2.0 * gc * st.rcim[p]);
        } else {
            mb[p] += denb[p] * 2.0 * st.mfit[p];
        }
    }
}

// ------------------------------------------------------------- k_grad
template <int NZ, bool MC>
__global__ void __launch_bounds__(kBlock) k_grad(const double* __restrict__ ev, int64_t ld, int64_t n, int scheme,
                                                const DevStatic* __restrict__ S, const DevDyn* __restrict__ D,
                                                const C64* __restrict__ cpl, GradLayout G, int NG, int NGp,
                                                double* __restrict__ gpart, Partial* __restrict__ part, int noint) {
    extern __shared__ double lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* stage = lds + wave * (kStage * 64);
    double* acc = lds + 4 * kStage * 64 + wave * NGp;
    int32_t* sidx = reinterpret_cast<int32_t*>(lds + 4 * kStage * 64 + 4 * NGp) + wave * kStage;
    for (int q = lane; q < NGp; q += 64) acc[q] = 0.0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    Emitter em{stage, sidx, acc, 0, lane};
    double s0 = 0, c0 = 0, s1 = 0, c1 = 0;

    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
        const int64_t gi = base + threadIdx.x;
        const bool valid = gi < n;
        const int64_t i = valid ? gi : n - 1;
        const double mkp = ld_d(ev, ld, i, E_MKP);
        const double mjp = ld_d(ev, ld, i, E_MJPSIP);
        const double mjk = ld_d(ev, ld, i, E_MJPSIK);
        const double w = ld_d(ev, ld, i, E_W);

        C64 H0[8];
        lambda_chain(S, D, cpl, ev, ld, i, mkp, H0);
        double fni = 0.0;
#pragma unroll
        for (int h = 0; h < 8; ++h) fni += norm2(H0[h]);
        if (S->nX > 0) {
            C64 HX[8];
            zcs_chain(S, D, cpl, ev, ld, i, mjk, HX);
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                fni += norm2(HX[h]);
                H0[h] = H0[h] + HX[h];
            }
        }
        double f = 0.0;
#pragma unroll
        for (int h = 0; h < 8; ++h) f += norm2(H0[h]);
        C64 Y[8];
#pragma unroll
        for (int h = 0; h < 8; ++h) Y[h] = H0[h];

        // Pc part: forward sums and the adjoint seeds
        ChainAng za;
        double zcp[8], zsp[8], cP = 0, sP = 0;
        C64 Sr[NZ > 0 ? NZ : 1][8], p0[NZ > 0 ? NZ : 1], Mo[NZ > 0 ? NZ : 1][NZ > 0 ? NZ : 1];
        double Md[NZ > 0 ? NZ : 1];
        C64 BW1[NZ > 0 ? NZ : 1], BW2[NZ > 0 ? NZ : 1][NZ > 0 ? NZ : 1];
        const int nb = S->scheme_nbins[scheme];
        const double sig = S->mres[0] + mjp * (S->mres[1] + mjp * S->mres[2]);
        if constexpr (NZ > 0) {
            za = load_chain(ev, ld, i, E_Z_CB);
            powers(ld_d(ev, ld, i, E_Z_CZ), zcp, 7);
            powers(ld_d(ev, ld, i, E_Z_SZ), zsp, 7);
            cP = ld_d(ev, ld, i, E_Z_CP);
            sP = ld_d(ev, ld, i, E_Z_SP);
#pragma unroll
            for (int r = 0; r < NZ; ++r) pc_angular(S, cpl, S->nR + r, za, zcp, zsp, cP, sP, Sr[r]);
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                C64 acc2;
                double d = 0.0;
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    cmac(acc2, conj(H0[h]), Sr[r][h]);
                    d += norm2(Sr[r][h]);
                }
                p0[r] = acc2;
                Md[r] = d;
#pragma unroll
                for (int q = 0; q < r; ++q) {
                    C64 mm;
#pragma unroll
                    for (int h = 0; h < 8; ++h) cmac(mm, conj(Sr[q][h]), Sr[r][h]);
                    Mo[q][r] = mm;
                }
                BW1[r] = C64(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < NZ; ++q) BW2[r][q] = C64(0.0, 0.0);
            }
            double accf = 0.0;
            C64 L[NZ];
            for (int b = 0; b < nb; ++b) {
                const double mb = mjp + S->scheme_z[scheme][b] * sig;
                pc_lineshapes<NZ>(S, D, mb, L);
                accf += pc_form<NZ>(L, p0, Md, Mo);
#pragma unroll
                for (int r = 0; r < NZ; ++r) {
                    BW1[r] = BW1[r] + L[r];
#pragma unroll
                    for (int q = 0; q < NZ; ++q) BW2[r][q] = BW2[r][q] + L[r] * conj(L[q]);
                }
            }
            const double inb = 1.0 / nb;
            f += accf * inb;
            if (MC && noint) {
                if (scheme != 0) pc_lineshapes<NZ>(S, D, mjp, L);
                C64 zero[NZ];
                fni += pc_form<NZ>(L, zero, Md, Mo);
            }
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                BW1[r] = BW1[r] * inb;
#pragma unroll
                for (int q = 0; q < NZ; ++q) BW2[r][q] = BW2[r][q] * inb;
#pragma unroll
                for (int h = 0; h < 8; ++h) cmac(Y[h], BW1[r], Sr[r][h]);
            }
        }

        // forward sums and the event's adjoint weight
        double v0, v1 = 0.0, aw;
        if (MC) {
            const bool unw = (w <= -1000.0);
            v0 = unw ? f : f * w;
            v1 = (noint && !unw) ? fni * w : 0.0;
            aw = unw ? 1.0 : w;
        } else {
            if (f > 0) v0 = log(f) * w;
            else v0 = (S->flags & PWA_F_NEG_FX_PENALTY) ? -1e12 : 0.0;
            v1 = (f > 0) ? 0.0 : 1.0;
            aw = (f > 0) ? w / f : 0.0;
        }
        if (!valid) {
            v0 = 0.0;
            v1 = 0.0;
            aw = 0.0;
        }
        kadd(s0, c0, v0);
        kadd(s1, c1, v1);
        const double a2 = 2.0 * aw;

        lambda_adjoint(S, D, cpl, ev, ld, i, mkp, Y, a2, G, em);
        if (S->nX > 0) zcs_adjoint(S, D, cpl, ev, ld, i, mjk, Y, a2, G, em);

        if constexpr (NZ > 0) {
            // Pc couplings: U_rh = BW1_r conj(H0_h) + sum_r' BW2_rr' conj(S_r'h)
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                C64 U[8];
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    C64 u = BW1[r] * conj(H0[h]);
#pragma unroll
                    for (int q = 0; q < NZ; ++q) u = u + BW2[r][q] * conj(Sr[q][h]);
                    U[h] = u;
                }
                pc_coupling_adjoint(S, S->nR + r, za, zcp, zsp, cP, sP, U, a2, G, em);
            }
            // Pc lineshapes: second pass over the sub-bins
            const double sc = a2 / nb;
            double gm[NZ], gw[NZ];
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                gm[r] = 0.0;
                gw[r] = 0.0;
            }
            const int ncls = S->km_nclass;
            for (int co = 0; co < (ncls > 0 ? ncls : 1); ++co) {
                const int ikc = (ncls > 0) ? S->km_class[co] : -1;
                double gb[16], fb[16], mbp[4];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    gb[q] = 0.0;
                    fb[q] = 0.0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) mbp[q] = 0.0;
                for (int b = 0; b < nb; ++b) {
                    const double mb = mjp + S->scheme_z[scheme][b] * sig;
                    C64 L[NZ], Mt[NZ], dL0[NZ], dL1[NZ];
                    double Vx[NZ];
                    KMState st;
                    C64 MtrC[8];
#pragma unroll
                    for (int r = 0; r < NZ; ++r) {
                        Mt[r] = C64(1.0, 0.0);
                        dL0[r] = C64(0.0, 0.0);
                        dL1[r] = C64(0.0, 0.0);
                    }
                    for (int c = 0; c < ncls; ++c) {
                        const int ik = S->km_class[c];
                        C64 Mtr[8];
                        if (ik == ikc) kmatrix_state(S, D, ik, mb, st, Mtr);
                        else kmatrix(S, D, ik, mb, Mtr);
                        if (ik == ikc) {
#pragma unroll
                            for (int q = 0; q < 8; ++q) MtrC[q] = Mtr[q];
                        }
#pragma unroll
                        for (int r = 0; r < NZ; ++r)
                            if (S->ik[S->nR + r] == ik) Mt[r] = select8(Mtr, S->ikm[S->nR + r]);
                    }
#pragma unroll
                    for (int r = 0; r < NZ; ++r) {
                        const int rr = S->nR + r;
                        const D2 m0 = D2::var(D->m0[rr], 0), g0 = D2::var(D->g0[rr], 1);
                        Cx<D2> Ls(D2(1.0), D2(0.0));
                        if (S->type[rr] == PWA_BREITWIGNER)
                            Ls = bw_twobody<D2>(S->P.m_p, S->P.m_jpsi, S->P.ffr, m0, g0, mb, S->L[rr]);
                        else
                            Ls = Cx<D2>(D2(Mt[r].re), D2(Mt[r].im));
                        const D2 Vv = vertex_pc<D2>(S->P, S->spin2[rr], mb, m0);
                        Vx[r] = Vv.v;
                        Ls = Ls * Vv;
                        L[r] = cval(Ls);
                        dL0[r] = cder(Ls, 0);
                        dL1[r] = cder(Ls, 1);
                    }
                    C64 Ab[8];
#pragma unroll
                    for (int r = 0; r < NZ; ++r) {
                        // P_r = p0_r + sum_q conj(L_q) M[q][r]
                        C64 P = p0[r];
                        rmac(P, Md[r], conj(L[r]));
#pragma unroll
                        for (int q = 0; q < NZ; ++q) {
                            if (q < r) P = P + conj(L[q]) * Mo[q][r];
                            else if (q > r) P = P + conj(L[q]) * conj(Mo[r][q]);
                        }
                        if (co == 0) {
                            gm[r] += sc * re_mul(P, dL0[r]);
                            gw[r] += sc * re_mul(P, dL1[r]);
                        }
                        const int rr = S->nR + r;
                        if (S->ik[rr] == ikc && ikc >= 0) {
                            // L = V * Mtr[ikm]: Abar[ikm] += sc * P * V
                            const C64 add = P * (Vx[r] * sc);
#pragma unroll
                            for (int q = 0; q < 8; ++q)
                                if (q == S->ikm[rr]) Ab[q] = Ab[q] + add;
                        }
                    }
                    if (ikc >= 0) kmatrix_adjoint(D, ikc, st, Ab, gb, fb, mbp);
                }
                if (ikc >= 0) {
                    const int o = G.km(ikc);
                    for (int q = 0; q < 16; ++q) em.emit(o + q, gb[q]);
                    for (int q = 0; q < 16; ++q) em.emit(o + 16 + q, fb[q]);
                    for (int q = 0; q < 4; ++q) em.emit(o + 32 + q, mbp[q]);
                }
            }
#pragma unroll
            for (int r = 0; r < NZ; ++r) {
                em.emit(G.res(S->nR + r), gm[r]);
                em.emit(G.res(S->nR + r) + 1, gw[r]);
            }
        }
    }
    if (em.cnt > 0) em.flush();
    block_reduce2(s0, c0, s1, c1);
    if (threadIdx.x == 0) part[blockIdx.x] = Partial{s0, c0, s1, c1};
    __syncthreads();
    for (int q = threadIdx.x; q < NG; q += kBlock) {
        double t = 0.0;
#pragma unroll
        for (int wv = 0; wv < 4; ++wv) t += lds[4 * kStage * 64 + wv * NGp + q];
        gpart[(int64_t)blockIdx.x * NG + q] = t;
    }
}

template <int NZ>
static hipError_t launch_grad_t(bool mc, int scheme, const double* ev, int64_t ld, int64_t n, const DevStatic* S,
                                const DevDyn* D, const C64* cpl, int ncpl, int nres, int NG, int blocks,
                                double* gpart, Partial* part, int noint, hipStream_t st) {
    GradLayout G{ncpl, nres};
    const int NGp = (NG + 1) & ~1;
    const size_t lds = sizeof(double) * (4 * kStage * 64 + 4 * NGp) + sizeof(int32_t) * 4 * kStage;
    if (mc) {
        hipError_t e = hipFuncSetAttribute((const void*)k_grad<NZ, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_grad<NZ, true>), dim3(blocks), dim3(kBlock), lds, st, ev, ld, n, scheme, S, D, cpl, G, NG,
                           NGp, gpart, part, noint);
    } else {
        hipError_t e = hipFuncSetAttribute((const void*)k_grad<NZ, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_grad<NZ, false>), dim3(blocks), dim3(kBlock), lds, st, ev, ld, n, scheme, S, D, cpl, G, NG,
                           NGp, gpart, part, noint);
    }
    return hipGetLastError();
}

size_t grad_lds_bytes(int NG) {
    const int NGp = (NG + 1) & ~1;
    return sizeof(double) * (4 * kStage * 64 + 4 * NGp) + sizeof(int32_t) * 4 * kStage;
}

hipError_t launch_grad(int nZ, bool mc, int scheme, const double* ev, int64_t ld, int64_t n, const DevStatic* S,
                       const DevDyn* D, const C64* cpl, int ncpl, int nres, int NG, int blocks, double* gpart,
                       Partial* part, int noint, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    switch (nZ) {
#define PWA_G(k) case k: return launch_grad_t<k>(mc, scheme, ev, ld, n, S, D, cpl, ncpl, nres, NG, blocks, gpart, part, noint, st);
    PWA_G(0) PWA_G(1) PWA_G(2) PWA_G(3) PWA_G(4) PWA_G(5) PWA_G(6) PWA_G(7) PWA_G(8)
#undef PWA_G
    default: return hipErrorInvalidValue;
    }
}

// out[k * NG + q] = sum over the partial rows [beg[k], beg[k+1]) of gpart[row * NG + q]
__global__ void __launch_bounds__(kBlock) k_grad_reduce(const double* __restrict__ gpart, const int32_t* __restrict__ beg,
                                                         int NG, double* __restrict__ out) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y;
    if (q >= NG) return;
    double s = 0.0, c = 0.0;
    for (int row = beg[k]; row < beg[k + 1]; ++row) kadd(s, c, gpart[(int64_t)row * NG + q]);
    out[k * NG + q] = s + c;
}

hipError_t launch_grad_reduce(const double* gpart, const int32_t* beg, int NG, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_grad_reduce, dim3((NG + kBlock - 1) / kBlock, 2), dim3(kBlock), 0, st, gpart, beg, NG, out);
    return hipGetLastError();
}

} // namespace pwa
