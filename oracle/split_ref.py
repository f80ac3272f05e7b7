"""Oracle: the Helmholtz split correction of list 1, numpy.

TEST INFRASTRUCTURE ONLY.  Restates, on the base quadrature (the
reference's default path; orders above 1 through ``extracted_terms`` and
``remainder(split_order=m)``):

* ``expansion_wrangler_fpnd.py:212-312`` -- ``_HelmholtzSplitSeriesRemainderKernel``:
  the smooth remainder ``G_k - G_0`` as a truncated series,
  2D ``c0 + sum_n s_n r^(2n) [-(1/2pi) log r + (H_n - log(k/2) - gamma)/(2pi) + i/4]``
  with ``s_n = (-1)^n (k^2/4)^n / (n!)^2``, ``c0 = i/4 - (log(k/2) + gamma)/(2pi)``;
  3D ``sum_{n>=1} (ik)^n r^(n-1) / (4 pi n!)``;
* ``4325-4389`` -- the series length rule (``_helmholtz_split_series_nmax``)
  from ``|k| * 2 sqrt(d) * (extent of the coarsest active box)``;
* ``5293-5317`` -- the self-pair limit (``c0`` in 2D, ``ik/(4 pi)`` in 3D),
  which the series also gives at r = 0;
* ``3491-3836`` -- the correction: a P2P sum over list 1 of
  ``remainder(x_t - y_s) * strength_s`` (strength = density x weight),
  self pairs excluded and replaced by ``limit * strength_t``;
* Yukawa: ``k = i lambda`` (``4175-4206``);
* axis target / source derivative wrappers around the Helmholtz kernel
  (``3884-3957``: the remainder P2P runs on the wrapped series kernel, its
  self limit is 0 for a wrapped chain, ``5293-5300``): the term-wise
  derivative ``d/dx_a R = (x_a - y_a) R'(r) / r`` (sign flipped for a
  source derivative).

Pins (tests): the series equals the closed form ``G_k - G_0`` from scipy's
Hankel function (2D) / ``exp`` (3D) where the series is well conditioned,
and its r -> 0 value equals the limit; the derivative series equals the
closed form ``R'(r)/r`` (``-(ik/4) H_1(kr) + 1/(2 pi r)`` over r in 2D,
``(e^{ikr}(ikr - 1) + 1)/(4 pi r^3)`` in 3D) and a central difference of
``remainder``.
"""

from __future__ import annotations

import math

import numpy as np

EULER_GAMMA = 0.5772156649015329


def series_nmax(dim, k, r_max, split_order=1):
    """``_helmholtz_split_series_nmax`` (4337-4389)."""
    if split_order < 1:
        raise ValueError("split_order must be >= 1")
    abs_k = float(abs(complex(k)))
    z_max = abs_k * float(r_max)
    tol, max_n = 1.0e-16, 96
    start_n = 1
    if z_max == 0.0:
        return start_n
    if dim == 3:
        mag = 1.0
        for i in range(1, start_n + 1):
            mag = mag * z_max / i
        n = start_n
        while n < max_n and mag > tol:
            n += 1
            mag = mag * z_max / n
        return min(max_n, n + 2)
    z2 = 0.25 * z_max * z_max
    mag = 1.0
    for i in range(1, start_n + 1):
        mag = mag * z2 / (i * i)
    log_r = abs(math.log(max(r_max, 1e-300)))
    log_k = abs(math.log(max(0.5 * abs_k, 1e-300)))
    scale = max(1.0, log_r, log_k + abs(EULER_GAMMA))
    n = start_n
    while n < max_n and mag * scale > tol:
        n += 1
        mag = mag * z2 / (n * n)
    return min(max_n, n + 2)


def max_neighbor_distance(tree):
    """``_helmholtz_split_max_neighbor_distance`` (4325-4335)."""
    lv = np.asarray(tree.box_levels)[np.asarray(tree.box_target_counts_nonchild) > 0]
    if lv.size == 0:
        return 0.0
    return float(2.0 * np.sqrt(tree.dim) * tree.root_extent * 0.5 ** int(lv.min()))


def extracted_terms(dim, k, split_order):
    """``_helmholtz_split_extra_terms`` (5052-5118): the non-smooth terms a
    split of order m takes out of the remainder into tables, as
    ``(kind, power, coefficient)``: 2D ``("power_log", 2n, -s_n / (2 pi))``
    for n < m; 3D ``("power", 2j - 1, (ik)^(2j) / (4 pi (2j)!))`` for j < m."""
    k = complex(k)
    out = []
    for j in range(1, int(split_order)):
        if dim == 2:
            s_n = ((-1) ** j) * (k * k / 4.0) ** j / (math.factorial(j) ** 2)
            out.append(("power_log", 2 * j, -s_n / (2.0 * np.pi)))
        else:
            n = 2 * j
            out.append(("power", n - 1, (1j * k) ** n / (4.0 * np.pi * math.factorial(n))))
    return out


def term_value(kind, power, r):
    """``r^p`` or ``r^p log r`` (0 at r = 0)."""
    r = np.asarray(r, dtype=np.float64)
    rp = r ** power
    if kind == "power":
        return rp
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(r <= 1e-300, 0.0, rp * np.log(np.where(r > 0, r, 1.0)))


def remainder(dim, k, nmax, r, split_order=1):
    """The series remainder kernel at distances ``r`` (any shape), complex.
    ``split_order`` m > 1 drops the terms ``extracted_terms`` names
    (expansion_wrangler_fpnd.py:284-304): 2D the log parts of n < m, 3D
    the even n <= 2(m - 1)."""
    k = complex(k)
    r = np.asarray(r, dtype=np.float64)
    out = np.zeros(r.shape, dtype=np.complex128)
    if dim == 2:
        if k == 0:
            return out
        lkh = np.log(0.5 * k)
        out += 0.25j - (1.0 / (2.0 * np.pi)) * (lkh + EULER_GAMMA)
        with np.errstate(divide="ignore", invalid="ignore"):
            logr = np.where(r <= 1e-300, 0.0, np.log(np.where(r > 0, r, 1.0)))
        for n in range(1, nmax + 1):
            s_n = ((-1) ** n) * (k * k / 4.0) ** n / (math.factorial(n) ** 2)
            h_n = float(np.sum(1.0 / np.arange(1, n + 1)))
            c_log = -s_n / (2.0 * np.pi)
            c_pow = s_n * ((h_n - (lkh + EULER_GAMMA)) / (2.0 * np.pi) + 0.25j)
            rp = r ** (2 * n)
            if n >= split_order:
                out += c_log * np.where(r <= 1e-300, 0.0, rp * logr)
            out += c_pow * rp
        return out
    for n in range(1, nmax + 1):
        if n % 2 == 0 and n <= 2 * (split_order - 1):
            continue
        out += (1j * k) ** n / (4.0 * np.pi * math.factorial(n)) * r ** (n - 1)
    return out


def remainder_radial_derivative(dim, k, nmax, r, split_order=1):
    """``R'(r) / r`` of the truncated series, term by term (0 at r = 0,
    where every derivative-wrapped pair value ``d_a R'(r)/r`` vanishes);
    ``split_order`` m > 1 drops the derivatives of the tabulated terms."""
    k = complex(k)
    r = np.asarray(r, dtype=np.float64)
    out = np.zeros(r.shape, dtype=np.complex128)
    pos = r > 1e-300
    rs = np.where(pos, r, 1.0)
    if dim == 2:
        if k == 0:
            return out
        lkh = np.log(0.5 * k)
        logr = np.log(rs)
        for n in range(1, nmax + 1):
            s_n = ((-1) ** n) * (k * k / 4.0) ** n / (math.factorial(n) ** 2)
            h_n = float(np.sum(1.0 / np.arange(1, n + 1)))
            c_log = -s_n / (2.0 * np.pi)
            c_pow = s_n * ((h_n - (lkh + EULER_GAMMA)) / (2.0 * np.pi) + 0.25j)
            # d/dr [c_log r^2n log r + c_pow r^2n] / r
            if n < split_order:
                c_log = 0.0
            out += rs ** (2 * n - 2) * (c_log * (2 * n * logr + 1.0) + 2 * n * c_pow)
    else:
        for n in range(2, nmax + 1):
            if n % 2 == 0 and n <= 2 * (split_order - 1):
                continue
            out += (n - 1) * (1j * k) ** n / (4.0 * np.pi * math.factorial(n)) * rs ** (n - 3)
    return np.where(pos, out, 0.0)


def closed_form_radial_derivative(dim, k, r):
    """``(G_k - G_0)'(r) / r`` from library functions (pins the above)."""
    r = np.asarray(r, dtype=np.float64)
    k = complex(k)
    if dim == 3:
        return (np.exp(1j * k * r) * (1j * k * r - 1.0) + 1.0) / (4.0 * np.pi * r ** 3)
    from scipy.special import hankel1

    return (-0.25j * k * hankel1(1, k * r) + 1.0 / (2.0 * np.pi * r)) / r


def self_limit(dim, k):
    """``_helmholtz_split_self_diagonal_limit`` (5293-5317)."""
    k = complex(k)
    if dim == 3:
        return 1j * k / (4.0 * np.pi)
    if k == 0:
        return 0.0j
    return 0.25j - (1.0 / (2.0 * np.pi)) * (np.log(0.5 * k) + EULER_GAMMA)


def closed_form_difference(dim, k, r):
    """``G_k(r) - G_0(r)`` from library functions (the pin of ``remainder``)."""
    r = np.asarray(r, dtype=np.float64)
    if dim == 3:
        return (np.exp(1j * k * r) - 1.0) / (4.0 * np.pi * r)
    from scipy.special import hankel1

    return 0.25j * hankel1(0, k * r) + np.log(r) / (2.0 * np.pi)


def _bary_interp(src, tgt):
    """Barycentric interpolation matrix (``_barycentric_interp_matrix``,
    120-141), weights prod_k 1/(x_j - x_k)."""
    w = np.array([1.0 / np.prod([xj - xk for xk in src if xk != xj]) for xj in src])
    out = np.empty((len(tgt), len(src)))
    for i, x in enumerate(tgt):
        d = x - np.asarray(src)
        if np.any(d == 0.0):
            out[i] = (d == 0.0).astype(float)
            continue
        out[i] = (w / d) / np.sum(w / d)
    return out


def smooth_sources(tree, f, m):
    """The smooth-quadrature sources of order m (5116-5275): per active box
    the m^d Gauss-Legendre nodes, strengths = f interpolated from the box's
    q^d nodes (tensor product, box-local node order from ``local_ids``) x
    extent^d x the Gauss weights.  Returns ``(points [dim, n_src], starts,
    counts, strengths [R, n_src])``."""
    dim, q = tree.dim, tree.quad_order
    f = np.atleast_2d(np.asarray(f, dtype=np.complex128))
    xq = 0.5 * (np.polynomial.legendre.leggauss(q)[0] + 1.0)
    xm, wm = np.polynomial.legendre.leggauss(m)
    xm, wm = 0.5 * (xm + 1.0), 0.5 * wm
    A = _bary_interp(xq, xm)
    grids = np.meshgrid(*([xm] * dim), indexing="ij")
    ref = np.stack([g.ravel() for g in grids])
    wref = np.prod(np.meshgrid(*([wm] * dim), indexing="ij"), axis=0).ravel()
    counts = np.asarray(tree.box_target_counts_nonchild)
    bstarts = np.asarray(tree.box_target_starts)
    starts = np.zeros(counts.size, np.int64)
    scounts = np.zeros(counts.size, np.int64)
    pts, st = [], []
    cur = 0
    for b in range(counts.size):
        starts[b] = cur
        if counts[b] == 0:
            continue
        h = tree.root_extent * 0.5 ** int(tree.box_levels[b])
        lo = np.asarray(tree.box_centers)[:, b] - 0.5 * h
        vals = np.empty((f.shape[0], q**dim), dtype=np.complex128)
        ids = np.asarray(tree.local_ids)[bstarts[b]:bstarts[b] + counts[b]]
        vals[:, ids] = f[:, bstarts[b]:bstarts[b] + counts[b]]
        v = vals.reshape((f.shape[0],) + (q,) * dim)
        if dim == 2:
            v = np.einsum("ai,bj,rij->rab", A, A, v)
        else:
            v = np.einsum("ai,bj,ck,rijk->rabc", A, A, A, v)
        pts.append(lo[:, None] + h * ref)
        st.append(v.reshape(f.shape[0], -1) * (h**dim * wref)[None, :])
        scounts[b] = m**dim
        cur += m**dim
    return np.concatenate(pts, axis=1), starts, scounts, np.concatenate(st, axis=1)


def correction(tree, strengths, k, nmax, target_boxes=None, starts=None, lists=None,
               deriv_axis=None, deriv_sign=1, split_order=1, kernel=None, sources=None,
               legacy=False, kernel_derivative=None):
    """P2P of the remainder over list 1: ``out[R, n_points]`` (complex);
    targets of unlisted boxes stay zero.  ``deriv_axis``: the axis of a
    target (``deriv_sign = +1``) or source (``-1``) derivative wrapper.
    ``kernel``: a radial function r -> complex used instead of the
    remainder (its value at r = 0 for self pairs).  ``sources``: ``(points,
    starts, counts)`` of a separate source set (``smooth_sources``); the
    strengths then index it and no pair is a self pair (r = 0 takes the
    series value at 0, which is the self limit).  ``legacy``: the closed
    form ``closed_form_difference`` off the diagonal and ``self_limit`` on
    it (the legacy subtraction, expansion_wrangler_fpnd.py:3806-3834).
    ``kernel_derivative``: with ``deriv_axis``, a radial ``r -> K'(r)/r``
    used instead of the remainder's (self pairs 0)."""
    st = np.atleast_2d(np.asarray(strengths, dtype=np.complex128))
    tb_list = tree.target_boxes if target_boxes is None else target_boxes
    starts = tree.neighbor_source_boxes_starts if starts is None else starts
    lists = tree.neighbor_source_boxes_lists if lists is None else lists
    out = np.zeros((st.shape[0], np.asarray(tree.points).shape[1]), dtype=np.complex128)
    pts = np.asarray(tree.points)
    bs, bc = tree.box_target_starts, tree.box_target_counts_nonchild
    for i, tb in enumerate(tb_list):
        t0, nt = int(bs[tb]), int(bc[tb])
        xt = pts[:, t0:t0 + nt]
        acc = np.zeros((st.shape[0], nt), dtype=np.complex128)
        for sb in lists[starts[i]:starts[i + 1]]:
            if sources is None:
                s0, ns = int(bs[sb]), int(bc[sb])
                ys = pts[:, s0:s0 + ns]
            else:
                s0, ns = int(sources[1][sb]), int(sources[2][sb])
                ys = np.asarray(sources[0])[:, s0:s0 + ns]
            dv = xt[:, :, None] - ys[:, None, :]  # [dim, nt, ns]
            r = np.sqrt(np.sum(dv ** 2, axis=0))
            same = (xt[:, :, None] == ys[:, None, :]).all(axis=0)
            if kernel is not None:
                rem = np.asarray(kernel(r), dtype=np.complex128)
            elif legacy and deriv_axis is None:
                rem = closed_form_difference(tree.dim, complex(k), np.where(same, 1.0, r))
                rem = np.where(same, self_limit(tree.dim, k), rem)
            elif deriv_axis is None:
                rem = remainder(tree.dim, k, nmax, r, split_order)
                rem = np.where(same, self_limit(tree.dim, k), rem)
            else:
                if kernel_derivative is not None:
                    rad = np.asarray(kernel_derivative(np.where(same, 1.0, r)), np.complex128)
                elif legacy:
                    rad = closed_form_radial_derivative(tree.dim, complex(k),
                                                        np.where(same, 1.0, r))
                else:
                    rad = remainder_radial_derivative(tree.dim, k, nmax, r, split_order)
                rem = deriv_sign * dv[deriv_axis] * rad
                rem = np.where(same, 0.0, rem)
            acc += st[:, s0:s0 + ns] @ rem.T
        out[:, t0:t0 + nt] = acc
    return out
