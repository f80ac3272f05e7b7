"""Pins of the Helmholtz split-correction oracle (CPU).

* the series remainder kernel equals the closed form G_k - G_0 (scipy's
  Hankel function in 2D, exp in 3D) wherever the series is well conditioned
  (|k| r_max <= 4), for real and Yukawa (k = i lambda) wave numbers;
* its r -> 0 value equals the self-pair limit;
* the series length rule stops when the term magnitude falls below 1e-16
  (closed-form check of the last term) and grows with |k| r_max;
* the list-1 P2P correction is linear in the strengths and, for one source
  box, equals the explicit double sum.
"""

import math

import numpy as np
import pytest

from oracle import split_ref as S
from volumential_amd import tree as vtree


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("k", [1.0, 5.0, 2.0j, 3.0 + 0.5j])
def test_series_matches_closed_form(dim, k):
    if dim == 2 and complex(k).imag != 0 and complex(k).real != 0:
        pytest.skip("scipy closed form checked for real and imaginary k")
    r_max = 4.0 / abs(k)
    nmax = S.series_nmax(dim, k, r_max)
    r = np.linspace(1e-3, r_max, 97)
    ser = S.remainder(dim, k, nmax, r)
    if dim == 2 and complex(k).real == 0:
        # Yukawa in 2D: G = K0(lam r)/(2 pi) = (i/4) H0(i lam r)
        from scipy.special import k0

        lam = complex(k).imag
        cf = k0(lam * r) / (2 * np.pi) + np.log(r) / (2 * np.pi)
    else:
        cf = S.closed_form_difference(dim, k, r)
    err = np.max(np.abs(ser - cf)) / max(1.0, np.max(np.abs(cf)))
    assert err <= 1e-13, err


@pytest.mark.parametrize("dim", [2, 3])
def test_series_at_zero_is_self_limit(dim):
    for k in (0.7, 12.0, 1.5j):
        nmax = S.series_nmax(dim, k, 1.0)
        assert abs(S.remainder(dim, k, nmax, np.array([0.0]))[0] - S.self_limit(dim, k)) < 1e-15
        # and the closed form approaches it
        if complex(k).imag == 0:
            r = 1e-7
            assert abs(S.closed_form_difference(dim, k, r) - S.self_limit(dim, k)) < 1e-6


@pytest.mark.parametrize("dim", [2, 3])
def test_series_nmax_rule(dim):
    n1 = S.series_nmax(dim, 1.0, 0.1)
    n2 = S.series_nmax(dim, 40.0, 0.35)
    assert 1 < n1 < n2 <= 96
    assert S.series_nmax(dim, 0.0, 1.0) == 1
    # the last accepted term magnitude is below tol (the rule adds 2 terms)
    z = 40.0 * 0.35
    n = n2 - 2
    if dim == 3:
        assert z**n / math.factorial(n) <= 1e-16
    else:
        assert (z * z / 4) ** n / math.factorial(n) ** 2 * max(1.0, abs(math.log(20.0)) + 0.5773) <= 1e-16


def test_correction_linear_and_explicit():
    tr = vtree.uniform_tree(2, 3, 2)
    k = 7.0
    nmax = S.series_nmax(2, k, S.max_neighbor_distance(tr))
    rng = np.random.default_rng(0)
    a = rng.standard_normal(tr.n_points) + 1j * rng.standard_normal(tr.n_points)
    b = rng.standard_normal(tr.n_points)
    ca = S.correction(tr, a, k, nmax)[0]
    cb = S.correction(tr, b, k, nmax)[0]
    cab = S.correction(tr, np.stack([a, 2 * a - 3 * b]), k, nmax)
    assert np.allclose(cab[0], ca, rtol=0, atol=1e-14)
    assert np.allclose(cab[1], 2 * ca - 3 * cb, rtol=0, atol=1e-13)
    # one target box, explicit double sum
    i = 5
    tb = tr.target_boxes[i]
    n = tr.n_q_points
    exp = np.zeros(n, complex)
    for sb in tr.neighbor_source_boxes_lists[tr.neighbor_source_boxes_starts[i]:
                                             tr.neighbor_source_boxes_starts[i + 1]]:
        for t in range(n):
            for s in range(n):
                xt = tr.points[:, tb * n + t]
                ys = tr.points[:, sb * n + s]
                r = float(np.linalg.norm(xt - ys))
                val = S.self_limit(2, k) if r == 0 else S.remainder(2, k, nmax, np.array([r]))[0]
                exp[t] += val * a[sb * n + s]
    assert np.allclose(ca[tb * n:(tb + 1) * n], exp, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("k", [1.0, 5.0, 2.0j])
def test_derivative_series_matches_closed_form(dim, k):
    """R'(r)/r of the series against the library closed form and against a
    central difference of the (already pinned) series itself."""
    r_max = 4.0 / abs(k)
    nmax = S.series_nmax(dim, k, r_max)
    # from 0.05 r_max on: below it the 3D closed form cancels (O(r^2) over r^3)
    r = np.linspace(0.05 * r_max, r_max, 61)
    ser = S.remainder_radial_derivative(dim, k, nmax, r)
    if dim == 2 and complex(k).real == 0:
        from scipy.special import k1

        lam = complex(k).imag
        # G = K0(lam r)/(2 pi): G' = -lam K1(lam r)/(2 pi); G_0' = -1/(2 pi r)
        cf = (-lam * k1(lam * r) / (2 * np.pi) + 1.0 / (2 * np.pi * r)) / r
    else:
        cf = S.closed_form_radial_derivative(dim, k, r)
    err = np.max(np.abs(ser - cf)) / max(1.0, np.max(np.abs(cf)))
    assert err <= 1e-12, err
    h = 1e-5
    fd = (S.remainder(dim, k, nmax, r + h) - S.remainder(dim, k, nmax, r - h)) / (2 * h) / r
    err = np.max(np.abs(ser - fd)) / max(1.0, np.max(np.abs(ser)))
    assert err <= 1e-7, err
    assert S.remainder_radial_derivative(dim, k, nmax, np.array([0.0]))[0] == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_derivative_correction_explicit(dim):
    """The derivative-wrapped correction is the explicit double sum of
    (x_a - y_a) R'(r)/r strength (self pairs 0); a source derivative is its
    negative."""
    tr = vtree.uniform_tree(dim, 2, 1)
    k = 3.0
    nmax = S.series_nmax(dim, k, S.max_neighbor_distance(tr))
    rng = np.random.default_rng(4)
    st = rng.standard_normal(tr.n_points) + 1j * rng.standard_normal(tr.n_points)
    axis = dim - 1
    got = S.correction(tr, st, k, nmax, deriv_axis=axis)[0]
    src = S.correction(tr, st, k, nmax, deriv_axis=axis, deriv_sign=-1)[0]
    assert np.array_equal(src, -got)
    pts = np.asarray(tr.points)
    bs, bc = tr.box_target_starts, tr.box_target_counts_nonchild
    want = np.zeros(tr.n_points, dtype=complex)
    for i, tb in enumerate(tr.target_boxes):
        for sb in tr.neighbor_source_boxes_lists[tr.neighbor_source_boxes_starts[i]:
                                                 tr.neighbor_source_boxes_starts[i + 1]]:
            for t in range(bs[tb], bs[tb] + bc[tb]):
                for s in range(bs[sb], bs[sb] + bc[sb]):
                    if s == t:
                        continue
                    d = pts[:, t] - pts[:, s]
                    r = np.linalg.norm(d)
                    want[t] += d[axis] * S.closed_form_radial_derivative(dim, k, r) * st[s]
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_series_roundoff_estimate_tracks_cancellation():
    """The 3D remainder (e^{ikr} - 1) / (4 pi r) summed as its power series in
    FP64 loses about unit roundoff times the largest term; the estimate the
    library warns with stays within two orders of magnitude of the measured
    loss (z = |k| r from 10 to 30, where the loss dominates)."""
    import cmath
    import math

    from volumential_amd.split import series_roundoff

    for z in (10.0, 18.0, 24.0, 30.0):
        term, acc = 1.0 + 0j, 0j
        for n in range(1, 200):
            term = term * 1j * z / n
            acc += term
        exact = cmath.exp(1j * z) - 1.0
        err = abs(acc - exact) / abs(exact)
        est = series_roundoff(3, z, 1.0)
        assert err / 100 <= est <= err * 100, (z, err, est)
        assert est > series_roundoff(3, z - 5.0, 1.0)
    assert series_roundoff(2, 40.0, 2 * math.sqrt(2) * 2 / 2**6) < 1e-15


@pytest.mark.parametrize("dim,k,m", [(2, 5.0, 2), (2, 3.0 + 1.0j, 3), (3, 4.0, 2), (3, 2.0j, 3),
                                     (3, 6.0, 4)])
def test_higher_order_remainder_plus_terms_is_order_one(dim, k, m):
    """A split of order m moves the terms ``extracted_terms`` names out of the
    series: remainder_m + sum coeff * term == remainder_1 at every r, and
    remainder_m equals the closed form G_k - G_0 minus those terms."""
    r = np.linspace(0.0, 0.6, 31)
    nmax = S.series_nmax(dim, k, 0.6)
    terms = S.extracted_terms(dim, k, m)
    assert len(terms) == m - 1
    assert all(t[0] == ("power_log" if dim == 2 else "power") for t in terms)
    assert [t[1] for t in terms] == ([2 * n for n in range(1, m)] if dim == 2 else
                                     [2 * j - 1 for j in range(1, m)])
    tab = sum(c * S.term_value(kind, p, r) for kind, p, c in terms)
    r1 = S.remainder(dim, k, nmax, r)
    rm = S.remainder(dim, k, nmax, r, split_order=m)
    assert np.allclose(rm + tab, r1, rtol=1e-13, atol=1e-15)
    rr = r[1:]
    cf = S.closed_form_difference(dim, k, rr) - sum(c * S.term_value(kind, p, rr)
                                                    for kind, p, c in terms)
    assert np.allclose(rm[1:], cf, rtol=1e-10, atol=1e-12)


def test_higher_order_extracted_coefficients():
    """The extracted terms are the series' own: 2D -s_n/(2 pi) on r^2n log r
    with s_n = (-1)^n (k^2/4)^n/(n!)^2; 3D (ik)^2j/(4 pi (2j)!) on r^(2j-1)."""
    k = 3.0
    (kind, p, c), = S.extracted_terms(2, k, 2)
    assert (kind, p) == ("power_log", 2) and np.isclose(c, (k * k / 4.0) / (2.0 * np.pi))
    (kind, p, c), (kind2, p2, c2) = S.extracted_terms(3, k, 3)
    assert (kind, p, kind2, p2) == ("power", 1, "power", 3)
    assert np.isclose(c, -(k * k) / (8.0 * np.pi)) and np.isclose(c2, k ** 4 / (96.0 * np.pi))


def test_split_terms_match_oracle_restatement():
    """The product's term list (split.split_terms) equals the oracle's."""
    from volumential_amd.split import split_terms

    for dim, k, m in ((2, 7.0, 4), (3, 2.5 + 0.5j, 4), (3, 1.0, 1)):
        a, b = split_terms(dim, k, m), S.extracted_terms(dim, k, m)
        assert [(x[0], x[1]) for x in a] == [(x[0], x[1]) for x in b]
        assert np.allclose([x[2] for x in a], [x[2] for x in b], rtol=1e-15)


def test_auto_split_order_ladder():
    """``helmholtz_split_order="auto"`` (expansion_wrangler_fpnd.py:4228-4330):
    rho = |k| h_max of the coarsest active box; default ladders 0.25 * 2^j
    (real part of k) and 0.5 * 2^j (imaginary part) over orders 2..12."""
    from volumential_amd.kernels import HelmholtzKernel
    from volumential_amd.split import choose_auto_split_order

    tr = vtree.uniform_tree(2, 3, 2)  # h_max = 0.5
    pick = lambda k, **cfg: choose_auto_split_order(tr, [HelmholtzKernel(2, k)], cfg)
    assert pick(0.4) == 2          # rho 0.2 <= 0.25
    assert pick(0.9) == 3          # rho 0.45 in (0.25, 0.5]
    assert pick(7.0) == 6          # rho 3.5 in (2, 4]
    assert pick(1e4) == 12         # beyond the ladder: order_max
    assert pick(1.2j) == 3         # imaginary: rho 0.6 in (0.5, 1]
    assert pick(0.4 + 1.2j) == 3   # the larger of the two picks
    assert pick(7.0, rho_thresholds=(0.5, 1.5, 3.0)) == 5
    assert pick(0.9, orders=(2, 4, 6, 8), rho_thresholds=(0.5, 1.5, 3.0)) == 2
    with pytest.raises(ValueError):
        pick(1.0, order_min=5, order_max=3)


@pytest.mark.parametrize("dim,q,m", [(2, 3, 5), (2, 4, 4), (3, 2, 4)])
def test_smooth_sources_integrate_the_interpolant(dim, q, m):
    """The smooth sources of order m carry the box interpolant: for f a
    polynomial of degree < q per axis their strengths sum, per box, to the
    box integral of f (= the q-point Gauss sum), and at m = q they are the
    base nodes with weight x f."""
    tr = vtree.uniform_tree(dim, q, 1)
    P = np.asarray(tr.points)
    f = (1.0 + 0.5 * P[0] - 0.25 * P[-1] ** (q - 1) + 0.1j * P[0] * P[-1] ** (q - 2))[None, :]
    pts, starts, counts, st = S.smooth_sources(tr, f, m)
    w = np.asarray(tr.weights)
    for b in np.flatnonzero(np.asarray(tr.box_target_counts_nonchild) > 0):
        s0, n = tr.box_target_starts[b], tr.box_target_counts_nonchild[b]
        base = np.sum(f[0, s0:s0 + n] * w[s0:s0 + n])
        assert np.isclose(np.sum(st[0, starts[b]:starts[b] + counts[b]]), base, rtol=1e-13)
        h = tr.root_extent * 0.5 ** tr.box_levels[b]
        c = np.asarray(tr.box_centers)[:, b]
        seg = pts[:, starts[b]:starts[b] + counts[b]]
        assert np.all(np.abs(seg - c[:, None]) < 0.5 * h)
    if m == q:
        assert np.allclose(np.sort(st[0].real), np.sort((f[0] * w).real), rtol=1e-13)


def test_smooth_correction_at_base_order_is_base_correction():
    """At m = q the smooth sources are the tree's own nodes: the correction
    over them (r = 0 taking the series value, the self limit) equals the
    base-quadrature correction."""
    tr = vtree.uniform_tree(2, 3, 2)
    P = np.asarray(tr.points)
    f = np.exp(-np.sum(P**2, axis=0))[None, :] * (1.0 + 0.3j)
    k = 5.0
    nmax = S.series_nmax(2, k, 1.5)
    base = S.correction(tr, f * np.asarray(tr.weights)[None, :], k, nmax)
    pts, starts, counts, st = S.smooth_sources(tr, f, 3)
    smooth = S.correction(tr, st, k, nmax, sources=(pts, starts, counts))
    assert np.allclose(smooth, base, rtol=1e-12, atol=1e-15)


def test_auto_smooth_quad_order_rule():
    """``"auto"`` smooth order (2649-2741): q + (m - 1) per order, plus
    ceil(rho_imag - 4) and ceil(0.5 (rho_real - 3)) above the hard
    thresholds, at least q, at most smooth_quad_order_max."""
    from volumential_amd.split import auto_smooth_quad_order as A

    assert A(6, 1, 0.5, 0.0) == 6
    assert A(6, 3, 0.5, 0.0) == 8
    assert A(6, 2, 5.0, 0.0) == 6 + 1 + 1          # real boost ceil(0.5 * 2)
    assert A(6, 2, 0.0, 6.5) == 6 + 1 + 3          # imaginary boost ceil(2.5)
    assert A(6, 4, 0.0, 6.5, {"smooth_quad_order_max": 9}) == 9
    assert A(6, 1, 0.0, 0.0, {"smooth_quad_order_min": 8}) == 8


@pytest.mark.parametrize("dim,k", [(2, 7.0), (2, 3.0 + 1.0j), (2, 4.0j), (3, 5.0), (3, 2.0 + 0.5j)])
def test_legacy_correction_equals_series_correction(dim, k):
    """The legacy subtraction (closed forms off the diagonal, the self limit
    on it; expansion_wrangler_fpnd.py:3806-3834) and the series path sum
    the same function G_k - G_0: equal to the series' accuracy.  A wrong
    sign or factor in either closed form, or a missing diagonal term,
    fails it."""
    tr = vtree.uniform_tree(dim, 3, 2)
    nmax = S.series_nmax(dim, k, S.max_neighbor_distance(tr))
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, tr.n_points)) + 1j * rng.standard_normal((2, tr.n_points))
    ser = S.correction(tr, a, k, nmax)
    leg = S.correction(tr, a, k, nmax, legacy=True)
    err = np.abs(leg - ser).max() / np.abs(ser).max()
    assert err < 1e-12


@pytest.mark.parametrize("dim,k,m", [(2, 7.0, 2), (2, 3.0 + 1.0j, 3), (3, 5.0, 2), (3, 2.0j, 3)])
def test_higher_order_derivative_plus_term_derivatives_is_order_one(dim, k, m):
    """Under a derivative wrapper: R_m'(r)/r plus each extracted term's
    coefficient times its radial derivative over r (2D r^2n log r ->
    r^(2n-2) (2n log r + 1), 3D r^p -> p r^(p-2)) is R_1'(r)/r; and the
    order-1 derivative matches the closed form."""
    r = np.linspace(0.05, 1.5, 31)
    nmax = 40
    tot = S.remainder_radial_derivative(dim, k, nmax, r, split_order=m)
    for kind, p, c in S.extracted_terms(dim, k, m):
        if kind == "power_log":
            tot = tot + c * r ** (p - 2) * (p * np.log(r) + 1.0)
        else:
            tot = tot + c * p * r ** (p - 2.0)
    one = S.remainder_radial_derivative(dim, k, nmax, r)
    np.testing.assert_allclose(tot, one, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(one, S.closed_form_radial_derivative(dim, k, r), rtol=1e-9,
                               atol=1e-12)


@pytest.mark.parametrize("dim,k", [(2, 4.0), (2, 2.0 + 1.0j), (3, 3.0), (3, 1.0 + 0.5j)])
def test_legacy_derivative_correction_equals_series_derivative(dim, k):
    """Under a derivative wrapper the legacy closed form d_a (G_k - G_0)'/r and
    the series derivative sum the same function (self pairs 0 in both)."""
    tr = vtree.uniform_tree(dim, 3, 2)
    nmax = S.series_nmax(dim, k, S.max_neighbor_distance(tr))
    rng = np.random.default_rng(5)
    a = rng.standard_normal((2, tr.n_points)) + 1j * rng.standard_normal((2, tr.n_points))
    for axis, sgn in ((0, 1), (dim - 1, -1)):
        ser = S.correction(tr, a, k, nmax, deriv_axis=axis, deriv_sign=sgn)
        leg = S.correction(tr, a, k, nmax, deriv_axis=axis, deriv_sign=sgn, legacy=True)
        assert np.abs(leg - ser).max() / np.abs(ser).max() < 1e-11
