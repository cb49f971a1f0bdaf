"""The polarization oracle (pnjl_polar_oracle.c) pinned without the
reference running: library quadrature, closed forms, principal-value
integrals by scipy, and exact invariants of the restated algorithm.

The reference's fixed rules carry ~1e-5 quadrature error, so checks against
exact integrals are loose; the exact invariants (linearity in xi, mirror and
k0 symmetries, channel difference, the xi = 0 reductions) are tight."""
import numpy as np
import pytest
from scipy import integrate

HB = 197.327
LAM = 602.3 / HB


@pytest.fixture(scope="module")
def O(oracle):
    oracle._polar_lib()
    return oracle


def f_plus(E, mu, T, Phi, Phib):
    e = np.clip(np.exp(-(E - mu) / T), 1e-200, 1e200)
    return (Phi * e + 2 * Phib * e * e + e ** 3) / (1 + 3 * Phi * e + 3 * Phib * e * e + e ** 3)


def f_minus(E, mu, T, Phi, Phib):
    e = np.clip(np.exp(-(E + mu) / T), 1e-200, 1e200)
    return (Phib * e + 2 * Phi * e * e + e ** 3) / (1 + 3 * Phib * e + 3 * Phi * e * e + e ** 3)


def test_quadrature_tables(O):
    t = O.polar_tables()
    x16, w16 = np.polynomial.legendre.leggauss(16)
    x32, w32 = np.polynomial.legendre.leggauss(32)
    assert np.max(np.abs(t[0:16] - x16)) < 1e-15 and np.max(np.abs(t[16:32] - w16)) < 1e-15
    assert np.max(np.abs(t[32:64] - x32)) < 1e-15 and np.max(np.abs(t[64:96] - w32)) < 1e-14
    pl_t16, pl_tp16 = t[96:112], t[112:128]
    # power map t = u^(1/0.35): its rule integrates smooth functions on [0, 1]
    for g, exact in ((lambda s: np.ones_like(s), 1.0), (lambda s: np.cos(s), np.sin(1.0))):
        assert abs(np.sum(w16 * pl_tp16 * g(pl_t16)) - exact) < 1e-6
    pr_t32, pr_tp32 = t[224:256], t[256:288]
    assert abs(np.sum(w32 * pr_tp32 * np.exp(pr_t32)) - (np.e - 1)) < 1e-8
    de_phi, de_phip = t[288:321], t[321:354]
    assert np.allclose(de_phi, -de_phi[::-1], atol=0) and abs(np.sum(de_phip) - 2.0) < 1e-6
    assert abs(0.5 * np.sum(de_phip * np.exp(de_phi)) - np.sinh(1.0)) < 1e-6


def test_const_term_and_A_against_quad(O):
    for m in (0.0, 1e-15, 0.02, 0.3, 1.7):
        num, _ = integrate.quad(lambda p: p * p / np.sqrt(p * p + m * m), 0, LAM, epsabs=1e-13, epsrel=1e-12)
        assert abs(O.A_const(m) - num) <= 1e-10 * max(1.0, num)
    xp, wp = np.polynomial.legendre.leggauss(64)
    xp, wp = 5 * (xp + 1), 5 * wp
    for m, mu, T, Phi, Phib in ((0.3, 0.1, 0.15, 0.4, 0.5), (1.5, 1.2, 0.6, 0.9, 0.95), (0.02, 0.0, 0.25, 0.0, 0.0)):
        dist, _ = integrate.quad(lambda p: p * p / np.sqrt(p * p + m * m) * (f_plus(np.sqrt(p * p + m * m), mu, T, Phi, Phib)
                                 + f_minus(np.sqrt(p * p + m * m), mu, T, Phi, Phib)), 0, 10, epsabs=1e-12, limit=200)
        expect = 4 * (-O.A_const(m) + dist)
        assert abs(O.A(m, mu, T, Phi, Phib, xp, wp) - expect) <= 5e-5 * abs(expect) + 1e-6


def test_B0_k0_against_principal_value_quad(O):
    m1, m2, mu, T, Phi, Phib = 0.3, 0.5, 0.1, 0.15, 0.3, 0.35
    for lam in (0.1, 0.25, 1.3, -0.7):
        for sg in (True, False):
            re, im = O.tilde_B0(sg, lam, 0.0, m1, m2, mu, T, Phi, Phib)
            Emax = np.sqrt(m1 ** 2 + LAM ** 2)
            d = (lam ** 2 + m1 ** 2 - m2 ** 2) / 2
            dist = (lambda E: f_plus(E, mu, T, Phi, Phib)) if sg else (lambda E: 1 - f_minus(E, mu, T, Phi, Phib))
            num = lambda E: np.sqrt(max(E * E - m1 * m1, 0.0)) * dist(E)
            E0 = -d / lam
            if m1 < E0 < Emax:
                pv, _ = integrate.quad(num, m1, Emax, weight="cauchy", wvar=E0, limit=400)
                exact = 2 * pv / lam
                assert abs(im - 2 * np.pi * num(E0) / lam) <= 1e-12 * abs(im)
            else:
                exact = 2 * integrate.quad(lambda E: num(E) / (lam * E + d), m1, Emax, limit=400)[0]
                assert im == 0.0
            assert abs(re - exact) <= 2e-5 * max(1.0, abs(exact)), (lam, sg, re, exact)


def test_B0_kpos_against_quad(O):
    m1, m2, mu, T, Phi, Phib = 0.3, 0.5, 0.1, 0.15, 0.3, 0.35
    Emax = np.sqrt(m1 ** 2 + LAM ** 2)
    for lam, k in ((0.9, 0.27), (0.3, 1.1), (-1.4, 0.6)):
        for sg in (True, False):
            re, im = O.tilde_B0(sg, lam, k, m1, m2, mu, T, Phi, Phib)
            dist = (lambda E: f_plus(E, mu, T, Phi, Phib)) if sg else (lambda E: 1 - f_minus(E, mu, T, Phi, Phib))

            def integrand(E):
                p = np.sqrt(max(E * E - m1 * m1, 0.0))
                c = (lam + E) ** 2 - m2 ** 2
                return dist(E) * np.log(abs((c - (p - k) ** 2) / (c - (p + k) ** 2)))
            # logarithmic singularities: split at the zeros of both arguments
            Es = np.linspace(m1, Emax, 20001)
            p = np.sqrt(np.maximum(Es ** 2 - m1 ** 2, 0))
            c = (lam + Es) ** 2 - m2 ** 2
            brk = []
            for g in (c - (p - k) ** 2, c - (p + k) ** 2):
                s = np.flatnonzero(np.sign(g[1:]) != np.sign(g[:-1]))
                brk += list(Es[s])
            pts = sorted(set([m1, Emax] + brk))
            exact = sum(integrate.quad(integrand, a, b, limit=400)[0] for a, b in zip(pts[:-1], pts[1:]) if b > a) / k
            assert abs(re - exact) <= 1e-3 * max(1.0, abs(exact)), (lam, k, sg, re, exact)
            # imaginary part: pi sign(lam)/k times the integral of the distribution over the cut
            cut = (np.abs(c - p * p - k * k) < 2 * p * k)
            ex_im = np.pi * np.sign(lam) / k * np.trapezoid(np.where(cut, dist(Es), 0.0), Es)
            assert abs(im - ex_im) <= 2e-3 * max(1.0, abs(ex_im)), (lam, k, sg, im, ex_im)


def test_B0_reference_unit_checks(O):
    """tests/unit/relaxtime/test_oneloopintegrals.jl:65-110 (mirror symmetry,
    k -> 0 continuity) and the lambda = 0, k = 0 NaN."""
    lam, m1, mu1, m2, mu2, T = 0.35, 0.28, 0.1, 0.51, 0.05, 0.17
    r = O.B0(lam, 0.0, m1, mu1, m2, mu2, T)
    mirrored = O.B0(-lam, 0.0, m2, mu2, m1, mu1, T)
    assert np.allclose(r, mirrored, rtol=1e-8, atol=1e-10)
    near = O.B0(lam, 1e-4, m1, mu1, m2, mu2, T)
    assert np.allclose(near, r, rtol=5e-3, atol=1e-4)
    assert np.isnan(O.B0(0.0, 0.0, m1, mu1, m1, mu1, T)[1])
    assert abs(O.B0(-0.05, 0.0, m1, mu1, m2, mu2, T, 0.3, 0.4)[1]) > 0


def test_B0_correction_invariants(O):
    args = (0.7, 0.4, 0.31, 0.52, 0.12, 0.12, 0.16, 0.45, 0.5)
    assert O.B0_correction(*args, 0.0) == (0.0, 0.0)
    for lam, k in ((0.7, 0.4), (0.7, 0.0), (-1.1, 0.9), (2.3, 0.05)):
        a = (lam, k) + args[2:]
        r1 = np.array(O.B0_correction(*a, 0.13))
        r2 = np.array(O.B0_correction(*a, 0.26))
        rm = np.array(O.B0_correction(*a, -0.13))
        assert np.allclose(r2, 2 * r1, rtol=1e-12, atol=1e-15)
        assert np.allclose(rm, -r1, rtol=1e-12, atol=1e-15)


def test_polarization_invariants(O):
    m1, m2, mu, T, Phi, Phib = 0.3, 0.62, 0.2, 0.15, 0.4, 0.45
    A1, A2 = -1.3, -0.9
    factor = -3 / (8 * np.pi ** 2)
    for xi in (0.0, 0.25):
        for k0, k in ((0.4, 0.3), (1.7, 0.0), (0.9, 1.2)):
            p = O.polarization(1, k0, k, m1, m2, mu, mu, T, Phi, Phib, xi, A1, A2, 0)
            s = O.polarization(2, k0, k, m1, m2, mu, mu, T, Phi, Phib, xi, A1, A2, 0)
            b = np.array(O.B0(k0, k, m1, mu, m2, mu, T, Phi, Phib))
            if xi:
                b = b + np.array(O.B0_correction(k0, k, m1, m2, mu, mu, T, Phi, Phib, xi))
            # S - P = factor * ((m1+m2)^2 - (m1-m2)^2) * B0
            assert np.allclose(np.subtract(s, p), factor * 4 * m1 * m2 * b, rtol=1e-11, atol=1e-14)
            # num_s = 1 restores k0 -> -k0 symmetry at mu1 = mu2
            a = O.polarization(1, k0, k, m1, m2, mu, mu, T, Phi, Phib, xi, A1, A2, 1)
            b_ = O.polarization(1, -k0, k, m1, m2, mu, mu, T, Phi, Phib, xi, A1, A2, 1)
            assert np.allclose(a, b_, rtol=1e-12, atol=1e-15)
            # width -> 0 reduces to polarization_aniso
            w = O.polarization(1, k0, k, m1, m2, mu, mu, T, Phi, Phib, xi, A1, A2, 0, gamma=0.0)
            assert np.allclose(w, p, rtol=0, atol=0)
    # |xi| <= 1e-12 is exactly isotropic
    a = O.polarization(2, 0.8, 0.5, m1, m2, mu, mu, T, Phi, Phib, 1e-13, A1, A2, 0)
    b = O.polarization(2, 0.8, 0.5, m1, m2, mu, mu, T, Phi, Phib, 0.0, A1, A2, 0)
    assert a == b
    with pytest.raises(ValueError):
        O.polarization(3, 0.8, 0.5, m1, m2, mu, mu, T, Phi, Phib, 0.0, A1, A2, 0)


def test_A_aniso_reductions(O):
    xp, wp = np.polynomial.legendre.leggauss(64)
    xp, wp = 5 * (xp + 1), 5 * wp
    xc, wc = np.polynomial.legendre.leggauss(8)
    m, mu, T, Phi, Phib = 0.3, 0.15, 0.17, 0.5, 0.55
    iso = O.A(m, mu, T, Phi, Phib, xp, wp)
    assert abs(O.A_aniso(m, mu, T, Phi, Phib, 0.0, xp, wp, xc, wc) - iso) <= 1e-13 * abs(iso)
    # A_aniso - A = A_correction + O(xi^2)
    err = [abs(O.A_aniso(m, mu, T, Phi, Phib, xi, xp, wp, xc, wc) - iso
               - O.A_correction(m, mu, T, Phi, Phib, xi, xp, wp)) for xi in (0.02, 0.01)]
    assert 3.2 < err[0] / err[1] < 4.8
