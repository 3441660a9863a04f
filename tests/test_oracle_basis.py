"""Pins the oracle's bond basis: real spherical harmonics against scipy and
the addition theorem, Chebyshev against its closed form, the radial/cutoff
product against the reference's own radial.cpp/cutoffs.cpp (oracle/_ref),
and every gradient against finite differences (ctests/test_y_grad.cpp:141-275,
ctests/test_radial.cpp)."""
import numpy as np
import pytest
from scipy.special import eval_legendre, sph_harm_y


def real_sh_scipy(v, lmax):
    """Real Y_lm without the Condon-Shortley phase, from scipy's complex
    Y_l^m (which carries the phase)."""
    r = np.linalg.norm(v)
    th = np.arccos(v[2] / r)
    ph = np.arctan2(v[1], v[0])
    out = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            mm = abs(m)
            c = sph_harm_y(l, mm, th, ph) * (-1) ** mm
            out.append(np.sqrt(2) * c.real if m > 0 else np.sqrt(2) * c.imag if m < 0 else c.real)
    return np.array(out)


POINTS = [np.array(p, dtype=float) for p in
          [(2.16, 7.12, -3.14), (0.3, -0.4, 1.2), (1.0, 0.0, 0.0), (0.0, 0.0, 2.0),
           (0.0, 0.0, -1.5), (-0.7, 0.2, 0.0), (1e-3, 2e-3, -3.0)]]


@pytest.mark.parametrize("p", range(len(POINTS)))
def test_real_ylm_matches_scipy(oracle, p):
    v = POINTS[p]
    Y = oracle.sph_harm(*v, 10)[0]
    np.testing.assert_allclose(Y, real_sh_scipy(v, 10), atol=1e-12)


def test_ylm_l1_is_y_z_x(oracle):
    # y_grad.cpp:718-736: l = 1 orders y, z, x with positive coefficients.
    v = np.array([0.3, -1.1, 0.7])
    Y = oracle.sph_harm(*v, 1)[0]
    c = np.sqrt(3 / (4 * np.pi)) / np.linalg.norm(v)
    np.testing.assert_allclose(Y[1:], c * np.array([v[1], v[2], v[0]]), rtol=1e-14)
    assert Y[0] == pytest.approx(1 / (2 * np.sqrt(np.pi)), rel=1e-15)


def test_ylm_addition_theorem(oracle):
    rng = np.random.default_rng(3)
    for _ in range(5):
        u, w = rng.normal(size=3), rng.normal(size=3)
        Yu, Yw = oracle.sph_harm(*u, 10)[0], oracle.sph_harm(*w, 10)[0]
        cosg = u @ w / np.linalg.norm(u) / np.linalg.norm(w)
        for l in range(11):
            sl = slice(l * l, (l + 1) ** 2)
            assert np.dot(Yu[sl], Yw[sl]) == pytest.approx(
                (2 * l + 1) / (4 * np.pi) * eval_legendre(l, cosg), abs=1e-12)


@pytest.mark.parametrize("p", range(len(POINTS)))
def test_ylm_gradient_fd(oracle, p):
    v = POINTS[p]
    Y = oracle.sph_harm(*v, 10)
    h = 1e-6 * max(1.0, np.linalg.norm(v))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd = (oracle.sph_harm(*(v + e), 10)[0] - oracle.sph_harm(*(v - e), 10)[0]) / (2 * h)
        np.testing.assert_allclose(Y[1 + k], fd, atol=1e-6 / min(1.0, np.linalg.norm(v)))


def test_chebyshev_closed_form(oracle):
    rc, N = 5.0, 12
    for r in np.linspace(0.05, 5.0, 23):
        v = np.array([r, 0.0, 0.0])
        g = oracle.radial(*v, rc, N, cutoff_function="quadratic")
        x = r / rc
        T = np.cos(np.arange(N) * np.arccos(x))
        np.testing.assert_allclose(g[0], T * (r - rc) ** 2, atol=1e-12)


def test_radial_gradient_fd(oracle):
    rng = np.random.default_rng(5)
    for kind in ("quadratic", "cosine"):
        for _ in range(4):
            v = rng.normal(size=3)
            v *= rng.uniform(0.5, 4.5) / np.linalg.norm(v)
            g = oracle.radial(*v, 5.0, 10, cutoff_function=kind)
            h = 1e-6
            for k in range(3):
                e = np.zeros(3)
                e[k] = h
                fd = (oracle.radial(*(v + e), 5.0, 10, cutoff_function=kind)[0]
                      - oracle.radial(*(v - e), 5.0, 10, cutoff_function=kind)[0]) / (2 * h)
                np.testing.assert_allclose(g[1 + k], fd, atol=1e-7)


def test_cutoffs_closed_form(oracle):
    rc = 4.5
    for r in (0.5, 2.0, 4.4):
        gq = oracle.radial(r, 0.0, 0.0, rc, 1, cutoff_function="quadratic")
        gc = oracle.radial(r, 0.0, 0.0, rc, 1, cutoff_function="cosine")
        assert gq[0, 0] == pytest.approx((r - rc) ** 2, rel=1e-15)
        assert gc[0, 0] == pytest.approx(0.5 * (np.cos(np.pi * r / rc) + 1), rel=1e-15)
    # Outside the cutoff the basis is zero (cutoffs.cpp:47-51).
    assert np.all(oracle.radial(rc + 0.1, 0, 0, rc, 4, cutoff_function="quadratic") == 0)


def test_radial_matches_reference_source(oracle):
    """Against the reference's radial.cpp + cutoffs.cpp compiled in
    oracle/_ref (built from the reference tree where it is present)."""
    L = oracle.refsrc()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(11)
    for kind, kname in ((0, "quadratic"), (1, "cosine")):
        for _ in range(20):
            v = rng.normal(size=3)
            v *= rng.uniform(0.1, 6.5) / np.linalg.norm(v)
            rc, N = 5.5, 12
            out = np.zeros(4 * N)
            r = float(np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
            L.refsrc_radial(*v, r, rc, N, 0.0, kind, out)
            mine = oracle.radial(*v, rc, N, cutoff_function=kname)
            np.testing.assert_array_equal(mine.reshape(-1), out)
        # Chebyshev alone, including its early return outside [r1, r2].
        for r in (0.0, 1.3, 5.5, 5.6):
            out = np.zeros(2 * 8)
            L.refsrc_chebyshev(r, 8, 0.0, 5.5, out)
            if r > 5.5:
                assert np.all(out == 0)
            else:
                x = r / 5.5
                np.testing.assert_allclose(out[:8], np.cos(np.arange(8) * np.arccos(x)), atol=1e-12)
