"""Pins of the oracle (CPU restatement) against the reference's own code and
against closed forms.

* Face fluxes, the boundary Riemann flux, the NSCBC outlet rates,
  fast_pow_posx, the RK3 weights: bit-identical to the reference's header code
  compiled from its sources (oracle/_ref, built by oracle/Makefile where the
  reference checkout exists).
* gr_func (IdealGasSolver.cpp:736-745 lives in a .cpp and is not compiled
  here): pinned by its closed-form properties.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

D = C.POINTER(C.c_double)


def dp(a):
    return a.ctypes.data_as(D)


def random_state(rng, n):
    """IdealGasState records: u[3], sp_vol, p, h."""
    s = np.empty((n, 6))
    s[:, 0:3] = rng.normal(scale=1.0, size=(n, 3))
    s[:, 3] = rng.uniform(0.5, 2.0, size=n)
    s[:, 4] = rng.uniform(50.0, 90.0, size=n)
    s[:, 5] = rng.uniform(200.0, 300.0, size=n)
    return s


def random_aux(rng, n):
    """AuxGasProp records: mu_total loc_total half_usq rho a_sgs sos fgr symh symp."""
    a = np.empty((n, 9))
    a[:, 0] = rng.uniform(1e-4, 1e-3, n)
    a[:, 1] = rng.uniform(1e-4, 1e-3, n)
    a[:, 2] = rng.uniform(0.0, 1.0, n)
    a[:, 3] = rng.uniform(0.5, 1.5, n)
    a[:, 4] = rng.uniform(0.0, 1e-4, n)
    a[:, 5] = rng.uniform(8.0, 12.0, n)
    a[:, 6] = rng.uniform(0.0, 0.5, n)
    a[:, 7] = rng.uniform(0.0, 1e-2, n)
    a[:, 8] = rng.uniform(0.0, 1e-2, n)
    # exercise the Ma clip min(fgr*Ma, 0.5) and the h_fax floor max(., 0) in both branches
    a[::7, 6] = 0.5
    a[::5, 7] = 1e3
    return a


# decided from the file's presence alone: collecting this module loads no library
needs_ref = pytest.mark.skipif(not os.path.exists(orc.REF_SO), reason="reference pin harness oracle/_ref not built")


@needs_ref
def test_skew_flux_bitwise_equal_to_reference():
    rng = np.random.default_rng(1)
    L, R = orc.lib(), orc.ref()
    for _ in range(2000):
        nc = rng.normal(size=6)
        s = random_state(rng, 2)
        f_ref, aux_ref = np.zeros(5), np.zeros(6)
        R.ref_flux_skew(dp(f_ref), dp(nc), dp(s[0]), dp(s[1]), dp(aux_ref))
        f, ua = np.zeros(5), np.zeros(3)
        iv, hs, pa = C.c_double(), C.c_double(), C.c_double()
        L.og_flux_skew(dp(f), dp(nc), dp(s[0]), dp(s[1]), dp(ua), C.cast(C.pointer(iv), D),
                       C.cast(C.pointer(hs), D), C.cast(C.pointer(pa), D))
        assert np.array_equal(f, f_ref)
        assert np.array_equal(ua, aux_ref[:3])
        assert (iv.value, hs.value, pa.value) == tuple(aux_ref[3:])


@needs_ref
def test_symmetric_flux_bitwise_equal_to_reference():
    rng = np.random.default_rng(2)
    L, R = orc.lib(), orc.ref()
    for _ in range(2000):
        nc = rng.normal(size=6)
        s = random_state(rng, 2)
        ua = rng.normal(size=3)
        iv, hs, pa = rng.uniform(0.5, 2), rng.uniform(200, 300), rng.uniform(50, 90)
        f_ref, f = np.zeros(5), np.zeros(5)
        R.ref_flux_symmetric(dp(f_ref), dp(nc), dp(ua), iv, hs, pa, dp(s[0]), dp(s[1]))
        L.og_flux_symmetric(dp(f), dp(nc), dp(ua), iv, hs, pa, dp(s[0]), dp(s[1]))
        assert np.array_equal(f, f_ref)


@needs_ref
def test_compact_flux_bitwise_equal_to_reference():
    rng = np.random.default_rng(3)
    L, R = orc.lib(), orc.ref()
    for _ in range(2000):
        fc = np.empty(5)
        un = rng.normal(size=3)
        fc[:3] = un / np.linalg.norm(un)
        fc[3] = rng.uniform(0.01, 1.0)
        fc[4] = rng.uniform(0.001, 0.1)
        s = random_state(rng, 2)
        a = random_aux(rng, 2)
        ua = 0.5 * (s[0, :3] + s[1, :3])
        iv, hs = 2.0 / (s[0, 3] + s[1, 3]), rng.uniform(200, 300)
        g0, g1 = rng.normal(size=9), rng.normal(size=9)
        f0 = rng.normal(size=5)
        f_ref, f = f0.copy(), f0.copy()
        R.ref_flux_compact(dp(f_ref), dp(fc), dp(s[0]), dp(s[1]), dp(a[0]), dp(a[1]), dp(ua), iv, hs, dp(g0), dp(g1))
        L.og_flux_compact(dp(f), dp(fc), dp(s[0]), dp(s[1]), dp(a[0]), dp(a[1]), dp(ua), iv, hs, dp(g0), dp(g1))
        assert np.array_equal(f, f_ref)


@needs_ref
def test_fast_pow_and_rk_weights_equal_to_reference():
    rng = np.random.default_rng(4)
    L, R = orc.lib(), orc.ref()
    for x in rng.uniform(0.2, 5.0, 500):
        assert L.og_fast_pow_posx(x, 0.75) == R.ref_fast_pow_posx(x, 0.75)
    w = np.zeros(9)
    R.ref_erk_wgt3(dp(w))
    wl = (C.c_double * 9).in_dll(L, "og_erk_wgt3")
    assert np.array_equal(np.array(list(wl)), w)


def test_rk_weights_are_ssp_rk3():
    """The weights reproduce Shu-Osher SSP-RK3 in the low-storage increment form:
    cumulative weights 1, (1/4, 1/4), (1/6, 1/6, 2/3)."""
    wl = np.array(list((C.c_double * 9).in_dll(orc.lib(), "og_erk_wgt3"))).reshape(3, 3)
    cum = np.cumsum(wl, axis=0)
    assert np.allclose(cum[0], [1, 0, 0])
    assert np.allclose(cum[1], [0.25, 0.25, 0])
    assert np.allclose(cum[2], [1 / 6, 1 / 6, 2 / 3])


def test_gr_func_closed_form_properties():
    """Logistic sensor 0.5/(1+exp(-2(x-nrm)/w)), w = 5e-3: midpoint 1/4 at x = nrm,
    limits 0 and 1/2, slope 1/(2w)*1/2*... = 50 at the midpoint, floor at x = 0."""
    g = orc.lib().og_gr_func
    nrm = 0.03
    assert g(nrm, nrm) == 0.25
    assert g(1.0, nrm) == 0.5
    assert g(0.0, nrm) == pytest.approx(0.5 / (1.0 + math.exp(12.0)), rel=1e-15)
    h = 1e-7
    slope = (g(nrm + h, nrm) - g(nrm - h, nrm)) / (2 * h)
    assert slope == pytest.approx(0.5 * 2.0 / 5.0e-3 / 4.0, rel=1e-6)
    xs = np.linspace(0, 0.1, 101)
    vals = [g(x, nrm) for x in xs]
    assert all(b >= a for a, b in zip(vals, vals[1:]))


def test_fast_pow_matches_pow():
    L = orc.lib()
    for x in np.linspace(0.3, 3.0, 50):
        assert L.og_fast_pow_posx(x, 0.75) == pytest.approx(x ** 0.75, rel=1e-14)


@needs_ref
def test_reference_svd_regularised_inverse():
    """calcSvd33 (SvdNew.hpp:24) with the thresholded inverse of calcCvGradCoeff
    (StaticSolver.cpp:5713-5756) equals the numpy SVD pseudo-inverse used by the
    dense operator restatement (oracle/operators_dense.py)."""
    from oracle.operators_dense import regularised_update

    rng = np.random.default_rng(5)
    R = orc.ref()
    for trial in range(300):
        A = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
        if trial % 3 == 0:  # force a clipped direction
            u, s, vt = np.linalg.svd(A)
            s[2] = 0.2
            A = u @ np.diag(s) @ vt
        Acm = np.ascontiguousarray(A.T.reshape(9))  # column major
        U, V, sig, err = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(1)
        R.ref_svd33(dp(Acm), dp(U), dp(V), dp(sig), dp(err))
        c = rng.normal(size=3)
        # reference arithmetic of StaticSolver.cpp:5713-5756 on the reference SVD
        sinv = np.where(sig < 0.45, 0.0, 1.0 / np.where(sig == 0, 1, sig))
        Um = U.reshape(3, 3).T  # U[(j)*3+i] = U_ij
        Vm = V.reshape(3, 3).T
        b = A @ c
        tmp = Um.T @ (c - b)
        ref_g = c + (Vm * sinv[None, :]) @ tmp
        ours = regularised_update(A, c[None, :], 0.45)[0]
        assert np.allclose(ours, ref_g, rtol=1e-12, atol=1e-14), (A, sig)


@needs_ref
def test_oracle_svd33_bitwise_equal_to_reference():
    """og_svd33 (the Sigma model's SVD) restates calcSvd33 (SvdNew.hpp:24-124):
    same sweeps, same convergence test, same rounding."""
    rng = np.random.default_rng(13)
    L, R = orc.lib(), orc.ref()
    for trial in range(500):
        A = rng.normal(size=(3, 3)) * 10.0 ** rng.uniform(-4, 4)
        if trial % 5 == 1:
            A[2] = 0.0  # a two-dimensional velocity gradient
        if trial % 5 == 2:
            A = A - A.T  # pure rotation
        if trial % 5 == 3:
            A = np.diag(rng.normal(size=3))
        Acm = np.ascontiguousarray(A.T.reshape(9))
        U, V, s = np.zeros(9), np.zeros(9), np.zeros(3)
        Ur, Vr, sr, err = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(1)
        it = L.og_svd33(dp(Acm), dp(U), dp(V), dp(s))
        itr = R.ref_svd33(dp(Acm), dp(Ur), dp(Vr), dp(sr), dp(err))
        assert it == itr
        assert np.array_equal(s, sr) and np.array_equal(U, Ur) and np.array_equal(V, Vr), A


def _riemann_states(rng, n):
    """Interior / boundary state pairs spanning every branch of the HLLC flux:
    subsonic both ways (contact left or right of the face) and supersonic
    inflow and outflow (normal Mach up to 3)."""
    out = []
    for k in range(n):
        nb = rng.normal(size=3) * rng.uniform(0.1, 2.0)
        un = nb / np.linalg.norm(nb)
        a, b = np.empty(6), np.empty(6)
        for s in (a, b):
            s[3] = rng.uniform(0.5, 2.0)          # sp_vol
            s[4] = rng.uniform(0.5, 2.0)          # p
            T = s[4] * s[3]                       # R = 1
            s[5] = 1.4 / 0.4 * T                  # h = gamma R T / (gamma - 1)
            c = math.sqrt(1.4 * s[4] * s[3])
            mach = rng.uniform(-3.0, 3.0) if k % 2 else rng.uniform(-0.8, 0.8)
            s[0:3] = rng.normal(scale=0.3 * c, size=3) + mach * c * un
        out.append((nb, a, b))
    return out


@needs_ref
def test_riemann_flux_bitwise_equal_to_reference():
    rng = np.random.default_rng(6)
    L, R = orc.lib(), orc.ref()
    for nb, a, b in _riemann_states(rng, 3000):
        f_ref, f = np.zeros(5), np.zeros(5)
        R.ref_flux_riemann(dp(f_ref), dp(nb), dp(a), dp(b), 1.4, 1.4)
        L.og_flux_riemann(dp(f), dp(nb), dp(a), dp(b), 1.4, 1.4)
        assert np.array_equal(f, f_ref)


def test_riemann_flux_closed_forms():
    # equal states: the physical flux (rho u.n, rho u u.n + p n, (rhoE + p) u.n) |n|;
    # supersonic outflow takes the inside state alone, supersonic inflow the outside one
    rng = np.random.default_rng(7)
    L = orc.lib()

    def phys(s, nb):
        A = np.linalg.norm(nb)
        n = nb / A
        rho, u, p, h = 1.0 / s[3], s[0:3], s[4], s[5]
        un = u @ n
        E = rho * (h + 0.5 * u @ u) - p
        return np.concatenate([[rho * un], rho * u * un + p * n, [(E + p) * un]]) * A

    for nb, a, b in _riemann_states(rng, 400):
        f = np.zeros(5)
        L.og_flux_riemann(dp(f), dp(nb), dp(a), dp(a), 1.4, 1.4)
        np.testing.assert_allclose(f, phys(a, nb), rtol=1e-12, atol=1e-12 * np.abs(phys(a, nb)).max())
        n = nb / np.linalg.norm(nb)
        ca = math.sqrt(1.4 * a[4] * a[3])
        cb = math.sqrt(1.4 * b[4] * b[3])
        a2, b2 = a.copy(), b.copy()
        a2[0:3] += (4 * max(ca, cb) - a2[0:3] @ n) * n    # both leave through the face at Mach > 2
        b2[0:3] += (4 * max(ca, cb) - b2[0:3] @ n) * n
        L.og_flux_riemann(dp(f), dp(nb), dp(a2), dp(b2), 1.4, 1.4)
        np.testing.assert_allclose(f, phys(a2, nb), rtol=1e-12, atol=1e-12 * np.abs(phys(a2, nb)).max())
        a2[0:3] -= 8 * max(ca, cb) * n                    # both enter
        b2[0:3] -= 8 * max(ca, cb) * n
        L.og_flux_riemann(dp(f), dp(nb), dp(a2), dp(b2), 1.4, 1.4)
        np.testing.assert_allclose(f, phys(b2, nb), rtol=1e-12, atol=1e-12 * np.abs(phys(b2, nb)).max())


@needs_ref
def test_nscbc_outlet_rhs_bit_identical_to_reference():
    # calcNscbcIgOutletRhs (BcCommon.hpp:419-470) on random faces: subsonic and
    # supersonic outflow (the Ma >= 1 branch), inflow clipped to un = 0, the
    # relaxed (sigma >= 0) and constant-pressure (sigma < 0) forms
    L, R = orc.lib(), orc.ref()
    rng = np.random.default_rng(31)
    for k in range(3000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        bc = np.array([rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), *rng.normal(scale=0.8, size=3)])
        cv = bc * (1.0 + 0.05 * rng.normal(size=5))
        if k % 5 == 0:
            bc[2:] = 2.5 * n  # supersonic outflow
        sigma = -1.0 if k % 7 == 0 else rng.uniform(0.0, 1.0)
        a, b = np.zeros(5), np.zeros(5)
        ua = L.og_nscbc_outlet_rhs(dp(a), dp(n), 0.03, dp(bc), dp(cv), 1.4, 0.7, sigma, 1.1)
        ub = R.ref_nscbc_outlet_rhs(dp(b), dp(n), 0.03, dp(bc), dp(cv), 1.4, 0.7, sigma, 1.1)
        assert ua == ub
        np.testing.assert_array_equal(a, b)


@needs_ref
def test_nscbc_mdot_rhs_bit_identical_to_reference():
    # calcNscbcIgOutletMdotWeakRhs (BcCommon.hpp:474-525) on random faces:
    # subsonic and supersonic outflow, inflow clipped to un = 0, zone mass
    # fluxes below, at and above the target
    L, R = orc.lib(), orc.ref()
    rng = np.random.default_rng(41)
    for k in range(3000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        bc = np.array([rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), *rng.normal(scale=0.8, size=3)])
        cv = bc * (1.0 + 0.05 * rng.normal(size=5))
        if k % 5 == 0:
            bc[2:] = 2.5 * n
        target = rng.uniform(0.1, 2.0)
        avg = target if k % 11 == 0 else target * rng.uniform(0.5, 1.5)
        a, b = np.zeros(5), np.zeros(5)
        T_relax = 0.37 + 0.1 * (k % 9)
        ua = L.og_nscbc_mdot_rhs(dp(a), dp(n), 0.03, dp(bc), dp(cv), 1.4, avg, target, T_relax)
        ub = R.ref_nscbc_mdot_rhs(dp(b), dp(n), 0.03, dp(bc), dp(cv), 1.4, avg, target, T_relax)
        assert ua == ub
        np.testing.assert_array_equal(a, b)


def test_nscbc_mdot_rhs_closed_forms():
    # at rest with the face state equal to its cell's only the incoming wave
    # acts, L1 = -c (rhou_avg - rhou_target) / T_relax: a zone passing less
    # than its target lowers the face pressure (dp/dt = -L1/2 < 0) and pulls
    # the face along n; at the target every rate vanishes. On a moving
    # subsonic face the rates equal the pressure outlet's whose L1 is the same
    # (sigma (1 - Ma^2) c / L (p - p_ref) = -c (rhou_avg - rhou_target) / T)
    L = orc.lib()
    n = np.array([0.0, 0.6, 0.8])
    gamma, rho, p, T = 1.4, 1.2, 1.05, 0.5
    c = math.sqrt(gamma * p / rho)
    bc = np.array([rho, p, 0.0, 0.0, 0.0])
    r = np.zeros(5)
    L.og_nscbc_mdot_rhs(dp(r), dp(n), 0.1, dp(bc), dp(bc.copy()), gamma, 0.8, 1.0, T)
    L1 = -c * (0.8 - 1.0) / T
    assert L1 > 0 and r[1] < 0
    np.testing.assert_allclose(r[1], -0.5 * L1, rtol=1e-14)
    np.testing.assert_allclose(r[0], -0.5 * L1 / c ** 2, rtol=1e-14)
    np.testing.assert_allclose(r[2:], 0.5 * L1 * n / c / rho, rtol=1e-14, atol=1e-300)
    L.og_nscbc_mdot_rhs(dp(r), dp(n), 0.1, dp(bc), dp(bc.copy()), gamma, 1.0, 1.0, T)
    assert np.all(r == 0.0)
    bc = np.array([rho, p, 0.1, 0.2, 0.3])
    cv = bc * np.array([1.01, 0.98, 1.1, 0.9, 1.05])
    un = bc[2:] @ n
    Ma = un / c
    avg, tgt, sigma, Lr = 0.9, 1.1, 0.4, 0.7
    p_ref = p + (avg - tgt) * Lr / (T * sigma * (1.0 - Ma * Ma))
    rm, rp = np.zeros(5), np.zeros(5)
    L.og_nscbc_mdot_rhs(dp(rm), dp(n), 0.1, dp(bc), dp(cv), gamma, avg, tgt, T)
    L.og_nscbc_outlet_rhs(dp(rp), dp(n), 0.1, dp(bc), dp(cv), gamma, Lr, sigma, p_ref)
    np.testing.assert_allclose(rm, rp, rtol=1e-12, atol=1e-12 * np.abs(rp).max())


def _eq_case(L_or_R, fn, case):
    tau, q, Tw = C.c_double(case["tau"]), C.c_double(0.0), C.c_double(case["Tw"])
    rc = getattr(L_or_R, fn)(C.byref(tau), C.byref(q), case["u1"], case["T1"], case["rho"], case["y1"], case["mu"],
                             0.7, 1004.5, C.byref(Tw), case["relax"], case["yp0"], case["tol"], case["guess"])
    return rc, tau.value, q.value, Tw.value


def _eq_cases(rng, n):
    # viscous sublayer to the log layer (y1+ ~ 0.5 .. 3e3), adiabatic walls with
    # and without a guess of their temperature, isothermal walls hotter and
    # colder than the cell, u1 = 0
    for k in range(n):
        rho, mu, T1 = rng.uniform(0.5, 2.0), 10.0 ** rng.uniform(-5.5, -4), rng.uniform(250.0, 400.0)
        y1 = 10.0 ** rng.uniform(-4, -2)
        u1 = 0.0 if k % 23 == 0 else 10.0 ** rng.uniform(-1, 2)
        guess = k % 3 != 0
        Tw = -(T1 + rng.uniform(0.0, 20.0)) if k % 4 != 3 else T1 + rng.uniform(-60.0, 60.0)
        yield dict(tau=mu * u1 / y1 * rng.uniform(0.5, 3.0), Tw=Tw, u1=u1, T1=T1, rho=rho, y1=y1, mu=mu,
                   relax=(0.7, 0.5)[k % 2], yp0=(1.0, 0.5)[k % 2], tol=(1e-4, 1e-6)[k % 2], guess=int(guess))


@needs_ref
def test_equilibrium_wall_model_bit_identical_to_reference():
    L, R = orc.lib(), orc.ref()
    rng = np.random.default_rng(43)
    for case in _eq_cases(rng, 600):
        assert _eq_case(L, "og_wm_eq_solve_qt", case) == _eq_case(R, "ref_wm_eq_solve_qt", case), case


def test_equilibrium_wall_model_closed_forms():
    # thin viscous layer (y1+ ~ 0.1, the damped eddy viscosity negligible) and
    # nearly constant properties: plane Couette flow. Stress mu u1 / y1; an
    # adiabatic wall recovers T1 + Pr u1^2 / (2 cp) (recovery factor Pr); an
    # isothermal wall takes q_w = k (T1 - T_w) / y1 + tau u1 / 2, k = cp mu / Pr
    L = orc.lib()
    rho, mu, T1, cp, Pr = 1.2, 1.8e-5, 300.0, 1004.5, 0.7
    y1, u1 = 1e-5, 5.0
    case = dict(tau=mu * u1 / y1, Tw=-T1, u1=u1, T1=T1, rho=rho, y1=y1, mu=mu, relax=0.7, yp0=1.0, tol=1e-11,
                guess=0)
    rc, tau, q, Tw = _eq_case(L, "og_wm_eq_solve_qt", case)
    assert rc == 0 and q == 0.0
    np.testing.assert_allclose(tau, mu * u1 / y1, rtol=2e-3)
    np.testing.assert_allclose(-Tw - T1, Pr * u1 ** 2 / (2 * cp), rtol=2e-2)
    case.update(Tw=T1 - 0.5)
    rc, tau, q, Tw = _eq_case(L, "og_wm_eq_solve_qt", case)
    assert rc == 0 and Tw == T1 - 0.5
    np.testing.assert_allclose(q, cp * mu / Pr * 0.5 / y1 + tau * u1 / 2, rtol=5e-3)
    case.update(u1=0.0, Tw=-T1)
    assert _eq_case(L, "og_wm_eq_solve_qt", case)[:2] == (0, 0.0)


def test_nscbc_outlet_rhs_closed_forms():
    # at rest with the face state equal to its cell's, only the incoming
    # characteristic acts: L1 = sigma c (p_bc - p_ref) / L, so dp_bc/dt =
    # -L1 / 2 and drho_bc/dt = -L1 / (2 c^2) (isentropic), and u_bc moves along
    # n by L1 / (2 rho c); with p_bc = p_ref the face is at rest in every rate
    L = orc.lib()
    n = np.array([0.0, 0.6, 0.8])
    gamma, sigma, Lr, rho, p, p_ref = 1.4, 0.3, 0.5, 1.2, 1.05, 1.0
    c = math.sqrt(gamma * p / rho)
    bc = np.array([rho, p, 0.0, 0.0, 0.0])
    r = np.zeros(5)
    L.og_nscbc_outlet_rhs(dp(r), dp(n), 0.1, dp(bc), dp(bc.copy()), gamma, Lr, sigma, p_ref)
    L1 = sigma * c / Lr * (p - p_ref)
    np.testing.assert_allclose(r[1], -0.5 * L1, rtol=1e-14)
    np.testing.assert_allclose(r[0], -0.5 * L1 / c ** 2, rtol=1e-14)
    np.testing.assert_allclose(r[2:], 0.5 * L1 * n / c / rho, rtol=1e-14, atol=1e-300)
    bc[1] = p_ref
    L.og_nscbc_outlet_rhs(dp(r), dp(n), 0.1, dp(bc), dp(bc.copy()), gamma, Lr, sigma, p_ref)
    assert np.all(r == 0.0)


def _wm_cases(rng, n):
    # speeds, distances and viscosities from the viscous sublayer to y+ ~ 1e4,
    # with the laminar guess (-1), the previous stress's u_tau and poor guesses
    for k in range(n):
        rho = rng.uniform(0.5, 2.0)
        mu = 10.0 ** rng.uniform(-6, -3)
        y1 = 10.0 ** rng.uniform(-4, -1)
        u1 = 0.0 if k % 17 == 0 else 10.0 ** rng.uniform(-3, 1)
        guess = (-1.0, 0.0, 10.0 ** rng.uniform(-4, 0))[k % 3]
        yield u1, y1, rho, mu, guess


@needs_ref
def test_wall_model_bit_identical_to_reference():
    L, R = orc.lib(), orc.ref()
    rng = np.random.default_rng(31)
    for case in _wm_cases(rng, 3000):
        assert L.og_wm_solve_tau(*case) == R.ref_wm_solve_tau(*case), case


def test_wall_model_closed_forms():
    # the solved u_tau reproduces the law of the wall: u1 = u_tau u+(y1 u_tau / nu)
    # with u+ = y+ + a1 y+^2 below y* = 23.3244 and ln(y+)/0.41 + 5.2 above (C1 at
    # y*: a1 = (1/(kappa y*) - 1) / (2 y*)); in the sublayer limit tau -> mu u1 / y1
    L = orc.lib()
    kappa, B, ystar = 0.41, 5.2, 23.3244
    a1 = 0.5 * (1.0 / (kappa * ystar) - 1.0) / ystar
    assert abs((ystar + a1 * ystar ** 2) - (math.log(ystar) / kappa + B)) < 2e-4  # continuous at y*
    assert abs((1 + 2 * a1 * ystar) - 1 / (kappa * ystar)) < 1e-14                # and its slope
    rho, mu, y1 = 1.2, 1.8e-5, 1e-3
    nu = mu / rho
    for yplus_true in (0.5, 5.0, 20.0, 30.0, 300.0, 3000.0):
        u_tau = yplus_true * nu / y1
        up = yplus_true + a1 * yplus_true ** 2 if yplus_true <= ystar else math.log(yplus_true) / kappa + B
        u1 = u_tau * up
        tau = L.og_wm_solve_tau(u1, y1, rho, mu, -1.0)
        np.testing.assert_allclose(math.sqrt(tau / rho), u_tau, rtol=2e-6)
    u1 = 1e-6  # y+ ~ 0.008: the quadratic term is a 3e-4 correction
    tau = L.og_wm_solve_tau(u1, y1, rho, mu, -1.0)
    np.testing.assert_allclose(tau, mu * u1 / y1, rtol=1e-3)


def _q_wall_cases(rng, n):
    # friction Reynolds numbers u_tau rho y1 / mu from 0 through the branch
    # point 8.28314 to ~1e5, wall hotter and colder than the cell
    for k in range(n):
        rho = rng.uniform(0.5, 2.0)
        mu = 10.0 ** rng.uniform(-6, -3)
        y1 = 10.0 ** rng.uniform(-4, -1)
        re = 0.0 if k % 13 == 0 else 10.0 ** rng.uniform(-2, 5)
        u_tau = re * mu / (rho * y1)
        T1 = rng.uniform(250.0, 400.0)
        yield T1, T1 + rng.uniform(-80.0, 80.0), rho * u_tau ** 2, rho, mu, mu / 0.7, 1004.5, y1


@needs_ref
def test_wall_heat_flux_bit_identical_to_reference():
    L, R = orc.lib(), orc.ref()
    rng = np.random.default_rng(37)
    for case in _q_wall_cases(rng, 3000):
        assert L.og_wm_q_wall(*case) == R.ref_wm_q_wall(*case), case


def test_wall_heat_flux_closed_forms():
    # no stress: Fourier conduction lambda (T1 - Tw) / y1 (lambda = loc cp);
    # continuous where T* changes form (Re = y* = 8.28314); far out the thermal
    # log law q = rho cp u_tau (T1 - Tw) / T+ with T+ = Pr_t/kappa ln Re + C_t,
    # Pr_t = 0.85, kappa = 0.41, C_t = 3.9 (the laminar conductivity negligible)
    L = orc.lib()
    T1, Tw, rho, mu, cp, y1 = 320.0, 290.0, 1.1, 1.8e-5, 1004.5, 2e-3
    loc = mu / 0.7
    np.testing.assert_allclose(L.og_wm_q_wall(T1, Tw, 0.0, rho, mu, loc, cp, y1), loc * cp * (T1 - Tw) / y1,
                               rtol=1e-15)
    assert L.og_wm_q_wall(T1, T1, 0.3, rho, mu, loc, cp, y1) == 0.0
    ystar = 8.28314

    def q_at(re, locp=loc):
        u_tau = re * mu / (rho * y1)
        return L.og_wm_q_wall(T1, Tw, rho * u_tau ** 2, rho, mu, locp, cp, y1)
    lo, hi = q_at(ystar), q_at(ystar * (1 + 1e-12))
    assert abs(hi - lo) <= 1e-9 * abs(lo)
    for re in (1e6, 1e8):
        u_tau = re * mu / (rho * y1)
        Tplus = 0.85 / 0.41 * math.log(re) + 3.9
        np.testing.assert_allclose(q_at(re, 0.0), rho * cp * u_tau * (T1 - Tw) / Tplus, rtol=30 * ystar / re * math.log(re))


@needs_ref
def test_nscbc_inlet_rhs_bit_identical_to_reference():
    L, R = orc.lib(), orc.ref()
    rng = np.random.default_rng(47)
    for k in range(3000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        bc = np.array([rng.uniform(0.8, 1.2), rng.uniform(0.8, 1.2), *(-rng.uniform(0.0, 0.6) * n)])
        cv = bc * (1.0 + 0.05 * rng.normal(size=5))
        if k % 9 == 0:
            cv[2:] = 3.0 * n * -1.0  # u0 + c < 0: nothing written
        a = np.full(5, 7.0)
        b = np.full(5, 7.0)
        ra = L.og_nscbc_inlet_rhs(dp(a), dp(n), 0.03, dp(bc), dp(cv), 1.4)
        rb = R.ref_nscbc_inlet_rhs(dp(b), dp(n), 0.03, dp(bc), dp(cv), 1.4)
        assert ra == rb
        np.testing.assert_array_equal(a, b)


def test_nscbc_inlet_rhs_closed_form():
    # with the face's normal speed equal to the cell's (no velocity gradient)
    # only the pressure gradient drives it: L5 = (u0 + c) dp/dn and, with mdot
    # and T fixed, d rho_bc / dt = gamma L5 / ((M gamma - 1) c^2), M = un / c
    L = orc.lib()
    n = np.array([0.0, -1.0, 0.0])
    gamma, rho, p = 1.4, 1.1, 0.9
    c = math.sqrt(gamma * p / rho)
    un = -0.2 * c
    bc = np.array([rho, p, 0.0, un * n[1], 0.0])
    for dpc in (0.0, 0.02, -0.03):
        cv = bc.copy()
        cv[1] = p - dpc
        r = np.zeros(5)
        assert L.og_nscbc_inlet_rhs(dp(r), dp(n), 0.05, dp(bc), dp(cv), gamma) == 0
        L5 = (un + c) * (dpc / 0.05)
        M = un / c
        np.testing.assert_allclose(r[0], gamma * L5 / ((M * gamma - 1.0) * c * c), rtol=1e-12, atol=1e-300)
        assert np.all(r[1:] == 0.0)
