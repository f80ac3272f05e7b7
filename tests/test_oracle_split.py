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
