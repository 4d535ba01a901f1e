"""Pin the oracle's restatement of the lineshape types that no shipped fit
uses (spline MODELINDEPENDENT, revised K-matrix REVISEDKMATRIX, triangle cusp
TRIANGLE; src/kernel_calEva.cu:133-155, :404-452) against independent
mathematics: scipy's not-a-knot cubic spline, exact reproduction of cubics,
the one-member and Lambda(1405) limits of the revised K-matrix against the
pinned Breit-Wigner, the shared K(s) of a one-channel K-matrix, and the cusp's
closed-form inner integral against numerical quadrature.  The library
rejects these types (PWA_E_MODEL); that is checked here too."""
import ctypes as C

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.interpolate import CubicSpline

import oracle
from pwa_amd import _lib, abi, models

M_CHIC1 = 3.51067  # src/cu_DPFPropogator.cu:15


def _lambda_range(d):
    return d.m_p + d.m_k - 0.002, d.m_lb - d.m_jpsi + 0.002


# ------------------------------------------------------------------ spline
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_spline_coefficients_match_scipy_not_a_knot(seed):
    rng = np.random.default_rng(seed)
    for n in (4, 5, 11, 24):
        x = 1.4 + np.cumsum(rng.uniform(0.02, 0.1, n))  # non-uniform, well separated
        y = rng.standard_normal(n)
        got = oracle.spline_coeffs(x, y)
        cs = CubicSpline(x, y, bc_type="not-a-knot")
        ref = np.stack([cs.c[3], cs.c[2], cs.c[1], cs.c[0]], axis=1)  # (y, b, c, d) per interval
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got[:-1] - ref)) < 1e-11 * scale
        # last knot: the last cubic re-expanded at x[n-1] (value, slope, curvature/2, d)
        h = x[-1] - x[-2]
        assert abs(got[-1, 0] - y[-1]) < 1e-14
        assert abs(got[-1, 1] - cs(x[-1], 1)) < 1e-10 * scale
        assert abs(got[-1, 2] - cs(x[-1], 2) / 2) < 1e-9 * scale
        assert abs(got[-1, 3] - ref[-1, 3]) < 1e-11 * scale
        assert h > 0


def _cubic(m, c):
    return c[0] + m * (c[1] + m * (c[2] + m * c[3]))


def test_lambda_spline_reproduces_cubics_times_vertex():
    """Not-a-knot splines reproduce any cubic exactly; the Lambda* spline
    lineshape is that cubic times BTerm(J, m, m0 = -0.5), zero outside
    [m_p + m_K - 2 MeV, m_Lb - m_J/psi + 2 MeV]."""
    rng = np.random.default_rng(4)
    d = models.base_desc()
    lo, hi = _lambda_range(d)
    kx = np.sort(rng.uniform(lo + 0.01, hi - 0.01, abi.PWA_NUMSP))
    cre, cim = rng.standard_normal(4), rng.standard_normal(4)
    kv = _cubic(kx, cre) + 1j * _cubic(kx, cim)
    for spin2 in (1, 3):
        m = models.assemble(models.base_desc(), [models.lambda_spline_res(rng, spin2, -1, kx, kv)])
        assert m.desc.n_params == len(m.params) == abi.PWA_NPAR_LAMBDA_SPLINE
        assert oracle.lib().orc_fit_param_count(oracle.dref(m.desc)) == len(m.params)
        for mm in list(rng.uniform(lo, hi, 20)) + [lo + 1e-9, hi - 1e-9]:
            bt = oracle.lib().orc_bterm(oracle.dref(m.desc), 0, spin2 / 2, mm, -0.5)
            want = (_cubic(mm, cre) + 1j * _cubic(mm, cim)) * bt
            got = oracle.lineshape(m.desc, m.params, 0, mm)
            assert abs(got - want) < 1e-11 * max(abs(want), 1.0), (spin2, mm)
        for mm in (lo - 1e-6, hi + 1e-6, 1.3):
            assert oracle.lineshape(m.desc, m.params, 0, mm) == 0


def test_pc_spline_reproduces_cubics_and_cuts_at_4p6():
    rng = np.random.default_rng(5)
    d = models.base_desc()
    lo = d.m_p + d.m_jpsi - 0.002
    kx = np.linspace(4.05, 4.55, abi.PWA_NUMSPZ)
    cre, cim = rng.standard_normal(4), rng.standard_normal(4)
    kv = _cubic(kx - 4.3, cre) + 1j * _cubic(kx - 4.3, cim)
    m = models.assemble(models.base_desc(), [models.lambda_res(rng, 1, -1, 1.52, 0.02),
                                             models.pc_spline_res(rng, 3, -1, kx, kv)])
    assert m.desc.n_params == len(m.params) == abi.PWA_NPAR_LAMBDA + abi.PWA_NPAR_PC_SPLINE
    for mm in list(rng.uniform(lo, 4.6, 20)) + [4.6]:
        bt = oracle.lib().orc_bterm(oracle.dref(m.desc), 1, 1.5, mm, -0.5)
        want = (_cubic(mm - 4.3, cre) + 1j * _cubic(mm - 4.3, cim)) * bt
        got = oracle.lineshape(m.desc, m.params, 1, mm)
        assert abs(got - want) < 1e-10 * max(abs(want), 1.0), mm
    for mm in (4.6 + 1e-9, 5.0, lo - 1e-6):
        assert oracle.lineshape(m.desc, m.params, 1, mm) == 0


def _events(n, seed):
    return _lib.generate_phsp(n, seed)


def _nll(m, d, mc):
    sw = np.ones(d.shape[1])
    return oracle.nll(m.desc, m.params, d, sw, models.window_schemes(d[1]), mc, np.ones(mc.shape[1]),
                      models.window_schemes(mc[1]))


def test_constant_lambda_spline_is_the_non_resonant_term():
    """Knots all (1, 0): the spline is 1 on the whole K- p range, so the
    resonance is the NONRESONANT one with m0 = -0.5 -- equal NLL."""
    rng = np.random.default_rng(6)
    d = models.base_desc()
    lo, hi = _lambda_range(d)
    kx = np.linspace(lo + 0.05, hi - 0.05, abi.PWA_NUMSP)
    other = models.lambda_res(rng, 3, -1, 1.52, 0.016)
    spl = models.lambda_spline_res(rng, 1, -1, kx, np.ones(abi.PWA_NUMSP, complex), sc=(0.7, 0.2))
    nr = models.lambda_res(rng, 1, -1, -0.5, 0.0, abi.PWA_NONRESONANT, sc=(0.7, 0.2))
    nr.params[:12] = spl.params[:12]
    pc = models.pc_res(rng, 1, -1, 4.44, 0.02)
    m_s = models.assemble(models.base_desc(), [other, spl, pc])
    m_n = models.assemble(models.base_desc(), [other, nr, pc])
    ev, mc = _events(300, 11), _events(600, 12)
    a, b = _nll(m_s, ev, mc), _nll(m_n, ev, mc)
    assert np.allclose(a["fx_data"], b["fx_data"], rtol=1e-12)
    assert abs(a["nll"] - b["nll"]) < 1e-10 * abs(b["nll"])


# -------------------------------------------------------- revised K-matrix
def _inv(z):
    return 1.0 / z


@pytest.mark.parametrize("chain", [abi.PWA_CHAIN_LAMBDA, abi.PWA_CHAIN_PC])
def test_revised_kmatrix_single_member_is_breit_wigner(chain):
    rng = np.random.default_rng(7)
    if chain == abi.PWA_CHAIN_LAMBDA:
        cases = [(1, -1, 1.52, 0.016), (3, -1, 1.69, 0.06), (5, 1, 1.82, 0.08), (7, 1, 2.1, 0.2)]
        mk = lambda j2, p, m0, g0, t: models.lambda_res(rng, j2, p, m0, g0, t)
        grid = np.linspace(1.44, 2.52, 9)
    else:
        cases = [(1, -1, 4.31, 0.01), (3, -1, 4.46, 0.02), (5, 1, 4.5, 0.05)]
        mk = lambda j2, p, m0, g0, t: models.pc_res(rng, j2, p, m0, g0, t)
        grid = np.linspace(4.04, 5.1, 9)
    idx = 0 if chain == abi.PWA_CHAIN_LAMBDA else 1
    for j2, p, m0, g0 in cases:
        rk = models.assemble(models.base_desc(), ([] if idx == 0 else [models.lambda_res(rng, 1, -1, 1.5, .01)])
                             + [mk(j2, p, m0, g0, abi.PWA_REVISEDKMATRIX)])
        bw = models.assemble(models.base_desc(), ([] if idx == 0 else [models.lambda_res(rng, 1, -1, 1.5, .01)])
                             + [mk(j2, p, m0, g0, abi.PWA_BREITWIGNER)])
        for mm in grid:
            a = oracle.lineshape(rk.desc, rk.params, idx, mm)
            b = oracle.lineshape(bw.desc, bw.params, idx, mm)
            assert abs(a - b) < 1e-12 * abs(b), (j2, p, mm)


def test_revised_kmatrix_lambda1405_single_member_is_two_channel_bw():
    """Below the K- p threshold (m0 = 1.405, S wave) both the revised
    K-matrix and BW_AMP use the K- p + Sigma pi width; with L = 0 their
    normalisations agree."""
    rng = np.random.default_rng(8)
    rk = models.assemble(models.base_desc(), [models.lambda_res(rng, 1, -1, 1.405, 0.05, abi.PWA_REVISEDKMATRIX)])
    bw = models.assemble(models.base_desc(), [models.lambda_res(rng, 1, -1, 1.405, 0.05)])
    assert oracle.build_amp(rk.desc, rk.params)["L"][0] == 0
    for mm in np.linspace(1.435, 2.5, 12):
        a = oracle.lineshape(rk.desc, rk.params, 0, mm)
        b = oracle.lineshape(bw.desc, bw.params, 0, mm)
        assert abs(a - b) < 1e-12 * abs(b), mm


def test_revised_kmatrix_members_share_one_k_of_s():
    """Members r of one J^P class: R_r = factor_r / (D_r (1 - i K(s))) for
    poles, factor / (1 - i K) for the non-resonant member, with
    K = sum_poles m_j Gamma_j(s) / (m_j^2 - s) + K_NR.  So -Im(1/R)/Re(1/R)
    is the same K for every member, Re(1/R_r) equals Re(1/B_r) of the
    member's own Breit-Wigner B_r, and with no non-resonant member
    K = sum_j -Im(1/B_j)/Re(1/B_j).  Members of another J^P class decouple."""
    rng = np.random.default_rng(9)
    poles = [(1.52, 0.016), (1.69, 0.06), (2.05, 0.15)]
    for with_nr in (False, True):
        res = [models.lambda_res(rng, 3, -1, m0, g0, abi.PWA_REVISEDKMATRIX) for m0, g0 in poles]
        if with_nr:
            res.append(models.lambda_res(rng, 3, -1, -1.0, 0.4, abi.PWA_REVISEDKMATRIX))
        res.append(models.lambda_res(rng, 3, 1, 1.9, 0.1, abi.PWA_REVISEDKMATRIX))  # other class
        m = models.assemble(models.base_desc(), res)
        bws = [models.assemble(models.base_desc(), [models.lambda_res(rng, 3, -1, m0, g0)]) for m0, g0 in poles]
        for mm in np.linspace(1.45, 2.5, 11):
            R = [oracle.lineshape(m.desc, m.params, r, mm) for r in range(len(res))]
            B = [oracle.lineshape(b.desc, b.params, 0, mm) for b in bws]
            ks = [-_inv(z).imag / _inv(z).real for z in R[:len(poles) + with_nr]]
            assert np.ptp(ks) < 1e-9 * max(1.0, np.max(np.abs(ks))), (with_nr, mm, ks)
            for r in range(len(poles)):
                assert abs(_inv(R[r]).real - _inv(B[r]).real) < 1e-9 * abs(_inv(B[r]).real)
            if not with_nr:
                w = sum(-_inv(b).imag / _inv(b).real for b in B)
                assert abs(ks[0] - w) < 1e-9 * max(1.0, abs(w))
            # the other-class member alone: a plain Breit-Wigner
            solo = models.assemble(models.base_desc(), [models.lambda_res(rng, 3, 1, 1.9, 0.1)])
            s = oracle.lineshape(solo.desc, solo.params, 0, mm)
            assert abs(R[-1] - s) < 1e-12 * abs(s)


def test_revised_kmatrix_pc_cross_terms_use_kp_masses():
    """Pc members couple through KMatrixTerm, which always takes the K- p
    masses (src/cu_DPFPropogator.cu:1070-1085): 1/R_r = 1/B_r - i D_r W_j / f_r
    with W_j the K- p Breit-Wigner width term of member j."""
    rng = np.random.default_rng(10)
    (m1, g1), (m2, g2) = (4.31, 0.01), (4.45, 0.03)
    lam = models.lambda_res(rng, 1, -1, 1.5, 0.01)
    m = models.assemble(models.base_desc(), [lam, models.pc_res(rng, 1, -1, m1, g1, abi.PWA_REVISEDKMATRIX),
                                             models.pc_res(rng, 1, -1, m2, g2, abi.PWA_REVISEDKMATRIX)])
    b1 = models.assemble(models.base_desc(), [lam, models.pc_res(rng, 1, -1, m1, g1)])
    L = oracle.build_amp(m.desc, m.params)["L"]
    out = np.zeros(2)
    for mm in np.linspace(4.05, 5.0, 9):
        R = oracle.lineshape(m.desc, m.params, 1, mm)
        B = oracle.lineshape(b1.desc, b1.params, 1, mm)
        oracle.lib().orc_bw_amp(oracle.dref(m.desc), 0, m2, g2, mm, int(L[2]), oracle.dp(out))
        bk = 1.0 / complex(out[0], out[1])
        w = -bk.imag / bk.real  # m2 Gamma2(s) / (m2^2 - s) with K- p masses
        f_over_d = 1.0 / _inv(B).real  # factor_r / D_r
        want = _inv(B) - 1j * w / f_over_d
        assert abs(_inv(R) - want) < 1e-9 * abs(want), mm


# ------------------------------------------------------------- triangle
def _bc(d, alpha, mpp, mr, g):
    m1 = complex(mr, -g / 2)
    b = (alpha * (d.m_lb ** 2 - d.m_k ** 2 - mpp ** 2) + d.m_p ** 2 - M_CHIC1 ** 2) / mpp ** 2 + 1 - 0.001j
    c = (alpha * (m1 ** 2 - d.m_p ** 2) + d.m_p ** 2 - alpha * (1 - alpha) * d.m_k ** 2) / mpp ** 2
    return b, c


def _inner_quad(d, alpha, mpp, mr, g):
    """(1/2) int_{alpha-1}^0 dy / (y^2 + b y + c) by adaptive quadrature."""
    b, c = _bc(d, alpha, mpp, mr, g)
    f = lambda y: 1.0 / (y * y + b * y + c)
    kw = dict(epsabs=1e-14, epsrel=1e-12, limit=200)
    return 0.5 * complex(quad(lambda y: f(y).real, alpha - 1, 0, **kw)[0],
                         quad(lambda y: f(y).imag, alpha - 1, 0, **kw)[0])


def _quirk(d, alpha, mpp, mr, g):
    b, c = _bc(d, alpha, mpp, mr, g)
    z = 4 * c - b * b
    return z.imag < 0 and abs(z.imag / z.real) <= 0.001


ALPHAS = (np.arange(100) + 0.5) / 100


@pytest.mark.parametrize("mr,g,mpp", [(4.2, 0.05, 4.3), (4.2, 0.05, 4.6), (4.46, 0.05, 4.45), (4.3, 0.2, 5.1),
                                      (4.4, 0.1, 4.1)])
def test_cusp_closed_form_matches_quadrature(mr, g, mpp):
    """Away from the branch cuts the closed form is the integral, and the
    amplitude is its 100-point midpoint rule over alpha divided by m^2."""
    d = models.base_desc()
    inner = [oracle.cusp_inner(d, a, mr, g, mpp) for a in ALPHAS]
    for a, v in zip(ALPHAS, inner):
        assert not _quirk(d, a, mpp, mr, g)
        q = _inner_quad(d, a, mpp, mr, g)
        assert abs(v - q) < 1e-9 * abs(q), a
    got = oracle.cusp(d, mr, g, mpp)
    want = sum(0.01 * v for v in inner) / mpp ** 2
    assert abs(got - want) < 1e-13 * abs(want)


@pytest.mark.parametrize("mr,g,mpp", [(4.46, 0.01, 4.0), (4.2, 0.05, 4.7), (4.46, 0.05, 5.1), (4.4, 0.1, 4.9)])
def test_cusp_closed_form_off_the_integral(mr, g, mpp):
    """Where the reference's closed form departs from the integral, it is by
    its complex functions' branches only: at alphas where 4c - b^2 has a
    small negative imaginary part (|Im/Re| <= 1e-3) cuCsqrtf returns the
    conjugate root (RooTDalitz/cuComplex.h:641-655) -- the value is then the
    closed form with that root, evaluated here with numpy's arctan -- and
    elsewhere a principal-branch arctan jump changes it by an integer
    multiple of pi / r, r = sqrt(4c - b^2)."""
    d = models.base_desc()
    n_off = 0
    for a in ALPHAS:
        v = oracle.cusp_inner(d, a, mr, g, mpp)
        b, c = _bc(d, a, mpp, mr, g)
        r = np.sqrt(4 * c - b * b)
        if _quirk(d, a, mpp, mr, g):
            rc = np.conj(r)
            want = (np.arctan(b / rc) - np.arctan((b + 2 * (a - 1)) / rc)) / rc
            assert abs(v - want) < 1e-10 * abs(want), a
            n_off += 1
            continue
        q = _inner_quad(d, a, mpp, mr, g)
        k = (v - q) * r / np.pi
        assert abs(k - round(k.real)) < 1e-6, (a, k)
        n_off += round(k.real) != 0
    assert n_off > 0


def test_triangle_pc_lineshape_is_cusp_times_vertex():
    rng = np.random.default_rng(12)
    m = models.assemble(models.base_desc(), [models.lambda_res(rng, 1, -1, 1.5, 0.01),
                                             models.pc_res(rng, 1, -1, 4.2, 0.05, abi.PWA_TRIANGLE)])
    for mm in (4.1, 4.3, 4.6):
        bt = oracle.lib().orc_bterm(oracle.dref(m.desc), 1, 0.5, mm, 4.2)
        want = oracle.cusp(m.desc, 4.2, 0.05, mm) * bt
        assert abs(oracle.lineshape(m.desc, m.params, 1, mm) - want) < 1e-14 * abs(want)


# --------------------------------------------------------------- library
@pytest.mark.parametrize("case", ["lambda_spline", "pc_spline", "revised_km", "pc_revised_km", "triangle"])
def test_library_layout_and_acceptance(case):
    """The library counts every type's parameters as the oracle does and
    accepts every type on the chains where the reference evaluates it
    (spline and revised K-matrix: Lambda* and Pc; triangle cusp: Pc); on a
    Zcs resonance it rejects them (PWA_E_MODEL)."""
    rng = np.random.default_rng(13)
    if case == "lambda_spline":
        r = models.lambda_spline_res(rng, 1, -1, np.linspace(1.45, 2.5, abi.PWA_NUMSP), np.ones(abi.PWA_NUMSP, complex))
    elif case == "pc_spline":
        r = models.pc_spline_res(rng, 1, -1, np.linspace(4.05, 4.55, abi.PWA_NUMSPZ), np.ones(abi.PWA_NUMSPZ, complex))
    elif case == "revised_km":
        r = models.lambda_res(rng, 1, -1, 1.52, 0.02, abi.PWA_REVISEDKMATRIX)
    elif case == "pc_revised_km":
        r = models.pc_res(rng, 1, -1, 4.45, 0.02, abi.PWA_REVISEDKMATRIX)
    else:
        r = models.pc_res(rng, 1, -1, 4.2, 0.05, abi.PWA_TRIANGLE)
    m = models.assemble(models.base_desc(), [models.lambda_res(rng, 3, -1, 1.52, 0.016), r])
    n = C.c_int32()
    assert _lib.lib().pwa_fit_param_count(C.byref(m.desc), C.byref(n)) == abi.PWA_OK
    assert n.value == len(m.params) == oracle.lib().orc_fit_param_count(oracle.dref(m.desc))
    rc = _lib.lib().pwa_helicity_couplings(C.byref(m.desc), None, None, C.byref(n))
    assert rc == abi.PWA_OK
    got = _lib.helicity_couplings(m.desc, m.params)
    ref = oracle.build_amp(m.desc, m.params)["vamp"]
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))
    # on a Zcs resonance they are rejected
    d = m.desc
    z = models.zcs_res(rng, 2, 1, 4.0, 0.1)
    m2 = models.assemble(models.base_desc(), [models.lambda_res(rng, 3, -1, 1.52, 0.016), z])
    m2.desc.type[1] = {"lambda_spline": abi.PWA_MODELINDEPENDENT, "pc_spline": abi.PWA_MODELINDEPENDENT,
                       "revised_km": abi.PWA_REVISEDKMATRIX, "pc_revised_km": abi.PWA_REVISEDKMATRIX,
                       "triangle": abi.PWA_TRIANGLE}[case]
    assert _lib.lib().pwa_helicity_couplings(C.byref(m2.desc), None, None, C.byref(n)) == abi.PWA_E_MODEL
    assert d.n_res == 2
