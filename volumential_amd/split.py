"""Helmholtz split correction of list 1 (``vl1_split_correction``).

With ``helmholtz_split`` the reference evaluates list 1 of a Helmholtz
potential as the Laplace-table evaluation plus an online point-to-point sum
of the smooth remainder ``G_k - G_0`` over list 1
(``expansion_wrangler_fpnd.py:3491-3836``; ``volume_fmm.py:1000-1026`` adds
the two).  This module holds the host side of the correction for split
order 1 on the base quadrature: the series length rule
(``_helmholtz_split_series_nmax``, 4325-4389), the device tree arrays, and
the launch.
"""

from __future__ import annotations

import logging
import math

import numpy as np
import torch

from . import _capi

EULER_GAMMA = 0.5772156649015329

logger = logging.getLogger(__name__)


def series_nmax(dim, k, r_max, split_order=1):
    """Series terms so that the first dropped term is below 1e-16 (+2):
    ``expansion_wrangler_fpnd.py:4337-4389``."""
    if split_order < 1:
        raise ValueError("split_order must be >= 1")
    abs_k = float(abs(complex(k)))
    z = abs_k * float(r_max)
    tol, max_n, n = 1.0e-16, 96, 1
    if z == 0.0:
        return 1
    if dim == 3:
        mag = z
        while n < max_n and mag > tol:
            n += 1
            mag = mag * z / n
        return min(max_n, n + 2)
    if dim != 2:
        raise NotImplementedError("split remainder series supports only 2D/3D")
    z2 = 0.25 * z * z
    mag = z2
    scale = max(1.0, abs(math.log(max(r_max, 1e-300))),
                abs(math.log(max(0.5 * abs_k, 1e-300))) + EULER_GAMMA)
    while n < max_n and mag * scale > tol:
        n += 1
        mag = mag * z2 / (n * n)
    return min(max_n, n + 2)


def series_roundoff(dim, k, r_max):
    """Relative rounding level of the remainder series at the largest list-1
    distance: unit roundoff times the largest term magnitude (the magnitude
    recursion of ``series_nmax``; the remainder itself is of order one in
    the same units).  At large |k| r the terms grow to about e^{|k| r} before
    they fall, and the alternating sum loses that many digits; the
    reference's series (expansion_wrangler_fpnd.py:212-312) behaves the
    same.  An estimate, not a bound."""
    z = float(abs(complex(k))) * float(r_max)
    if z == 0.0:
        return 0.0
    big, n = 1.0, 1
    if dim == 3:
        mag = z
        while n < 96:
            big = max(big, mag)
            n += 1
            mag = mag * z / n
    else:
        z2 = 0.25 * z * z
        mag = z2
        while n < 96:
            big = max(big, mag)
            n += 1
            mag = mag * z2 / (n * n)
    return 1.1e-16 * big


def max_neighbor_distance(tree):
    """``2 sqrt(d)`` times the extent of the coarsest box holding points
    (``_helmholtz_split_max_neighbor_distance``, 4325-4335)."""
    lv = np.asarray(tree.box_levels)[np.asarray(tree.box_target_counts_nonchild) > 0]
    if lv.size == 0:
        return 0.0
    return float(2.0 * np.sqrt(tree.dim) * tree.root_extent * 0.5 ** int(lv.min()))


def self_limit(dim, k):
    """The r -> 0 value of the remainder (5293-5317)."""
    k = complex(k)
    if dim == 3:
        return 1j * k / (4.0 * np.pi)
    if k == 0:
        return 0j
    return 0.25j - (np.log(0.5 * k) + EULER_GAMMA) / (2.0 * np.pi)


def split_terms(dim, k, split_order):
    """The non-smooth terms a split of order m tabulates, as ``(kind, power,
    coefficient)`` (``_helmholtz_split_extra_terms``,
    expansion_wrangler_fpnd.py:5052-5118): 2D ``("power_log", 2n,
    -s_n / (2 pi))``, ``s_n = (-1)^n (k^2/4)^n / (n!)^2``, and 3D
    ``("power", 2j - 1, (ik)^(2j) / (4 pi (2j)!))``, for n, j = 1..m-1."""
    k = complex(k)
    out = []
    for j in range(1, int(split_order)):
        if dim == 2:
            s_n = ((-1) ** j) * (k * k / 4.0) ** j / (math.factorial(j) ** 2)
            out.append(("power_log", 2 * j, complex(-s_n / (2.0 * math.pi))))
        elif dim == 3:
            n = 2 * j
            out.append(("power", n - 1, complex((1j * k) ** n / (4.0 * math.pi * math.factorial(n)))))
        else:
            raise NotImplementedError("split terms exist only in 2D and 3D")
    return out


def _split_order_ladder(rho, thresholds, orders):
    """``_select_split_order_from_rho`` (385-393)."""
    for threshold, order in zip(thresholds, orders):
        if rho <= float(threshold):
            return int(order)
    return int(orders[-1])


def choose_auto_split_order(tree, kernels, config=None):
    """``helmholtz_split_order="auto"``: ``_choose_auto_helmholtz_split_order``
    (4228-4330).  rho = |Re k| h and |Im k| h for the coarsest active box
    extent h; the order is the larger pick of two ladders, by default
    ``0.25 * 2^j`` (real) and ``0.5 * 2^j`` (imaginary) for orders
    ``order_min .. order_max`` (2..12)."""
    cfg = dict(config or {})
    order_min, order_max = int(cfg.get("order_min", 2)), int(cfg.get("order_max", 12))
    if order_max < order_min:
        raise ValueError("order_max must be >= order_min")
    if "rho_thresholds" in cfg or "orders" in cfg:
        th_re = tuple(cfg.get("rho_thresholds", (0.5, 1.5, 3.0)))
        th_im = th_re
        orders = tuple(cfg["orders"]) if "orders" in cfg else tuple(
            range(order_min, order_min + len(th_re) + 1))
    else:
        b_re, b_im = float(cfg.get("rho_base_real", 0.25)), float(cfg.get("rho_base_imag", 0.5))
        if b_re <= 0.0 or b_im <= 0.0:
            raise ValueError("rho_base_real and rho_base_imag must be positive")
        count = max(0, order_max - order_min)
        th_re = tuple(b_re * 2.0 ** j for j in range(count))
        th_im = tuple(b_im * 2.0 ** j for j in range(count))
        orders = tuple(range(order_min, order_max + 1))
    th_re = tuple(cfg.get("rho_thresholds_real", th_re))
    th_im = tuple(cfg.get("rho_thresholds_imag", th_im))
    if "orders" not in cfg:
        if len(th_re) != len(th_im):
            raise ValueError("rho_thresholds_real and rho_thresholds_imag must have the same "
                             "length when orders are not provided")
        orders = tuple(range(order_min, order_min + len(th_re) + 1))
    if len(orders) != len(th_re) + 1:
        raise ValueError("orders must have one more element than thresholds")
    rho_re, rho_im = split_rho(tree, kernels)
    pick = max(_split_order_ladder(rho_re, th_re, orders), _split_order_ladder(rho_im, th_im, orders))
    if "orders" not in cfg:
        pick = max(order_min, min(order_max, pick))
    return pick


def interp_matrix(src_nodes, tgt_nodes):
    """Barycentric interpolation matrix from ``src_nodes`` to ``tgt_nodes``
    (exact at coinciding nodes; ``_barycentric_interp_matrix``, 120-141)."""
    from .quadrature import barycentric_weights

    src = np.asarray(src_nodes, np.float64)
    tgt = np.asarray(tgt_nodes, np.float64)
    if src.size == 1:
        return np.ones((tgt.size, 1))
    w = barycentric_weights(src)
    out = np.empty((tgt.size, src.size))
    for i, x in enumerate(tgt):
        d = x - src
        hit = np.flatnonzero(d == 0.0)
        if hit.size:
            out[i] = 0.0
            out[i, hit[0]] = 1.0
            continue
        t = w / d
        out[i] = t / t.sum()
    return out


class SmoothSources:
    """Sources of the smooth-quadrature correction path (5116-5275): per
    active box the m^d tensor Gauss-Legendre nodes of order m (the smooth
    quadrature order), strengths = the density interpolated from the box's
    q^d nodes times extent^d times the Gauss weights.  Geometry is built
    once per (tree, m); ``strengths`` interpolates on the GPU."""

    def __init__(self, tree, m):
        from .quadrature import gauss_legendre_01

        dim, q = int(tree.dim), int(tree.quad_order)
        self.dim, self.q, self.m = dim, q, int(m)
        xq, _ = gauss_legendre_01(q)
        xm, wm = gauss_legendre_01(self.m)
        self.interp = interp_matrix(xq, xm)  # [m, q]
        grids = np.meshgrid(*([xm] * dim), indexing="ij")
        ref = np.stack([g.ravel() for g in grids])  # [dim, m^d]
        wref = np.prod(np.meshgrid(*([wm] * dim), indexing="ij"), axis=0).ravel()
        counts = np.asarray(tree.box_target_counts_nonchild, np.int64)
        starts = np.asarray(tree.box_target_starts, np.int64)
        levels = np.asarray(tree.box_levels, np.int64)
        centers = np.asarray(tree.box_centers, np.float64)
        ids = np.asarray(tree.local_ids, np.int64)
        nq, nm = q**dim, self.m**dim
        active = np.flatnonzero(counts > 0)
        if np.any(counts[active] != nq):
            raise ValueError(f"the smooth correction requires exactly {nq} sources per active box")
        sstarts = np.zeros(counts.size, np.int32)
        scounts = np.zeros(counts.size, np.int32)
        pts = np.empty((dim, active.size * nm))
        sw = np.empty(active.size * nm)
        gather = np.empty((active.size, nq), np.int64)  # point of each box mode
        for i, b in enumerate(active):
            sstarts[b], scounts[b] = i * nm, nm
            h = float(tree.root_extent) * 0.5 ** int(levels[b])
            lo = centers[:, b] - 0.5 * h
            pts[:, i * nm:(i + 1) * nm] = lo[:, None] + h * ref
            sw[i * nm:(i + 1) * nm] = h**dim * wref
            gather[i, ids[starts[b]:starts[b] + nq]] = np.arange(starts[b], starts[b] + nq)
        dev = "cuda"
        self.n_src = int(active.size * nm)
        self.points = torch.as_tensor(pts, device=dev)
        self.box_starts = torch.as_tensor(sstarts, device=dev)
        self.box_counts = torch.as_tensor(scounts, device=dev)
        self.weights = torch.as_tensor(sw, dtype=torch.complex128, device=dev)
        self._gather = torch.as_tensor(gather.ravel(), device=dev)
        self._n_active = int(active.size)
        self._interp = torch.as_tensor(self.interp, dtype=torch.complex128, device=dev)
        self._box_of_source = np.repeat(active, nm)

    def strengths(self, f):
        """``f[R, n_points]`` (complex, box-node values) -> ``[R, n_src]``."""
        R = int(f.shape[0])
        dim, q = self.dim, self.q
        v = f[:, self._gather].reshape((R, self._n_active) + (q,) * dim)
        A = self._interp
        if dim == 2:
            v = torch.matmul(torch.matmul(A, v), A.T)
        else:
            v = torch.einsum("ai,rbijk->rbajk", A, v)
            v = torch.einsum("bj,rnajk->rnabk", A, v)
            v = torch.einsum("ck,rnabk->rnabc", A, v)
        return v.reshape(R, self.n_src) * self.weights[None, :]

    def per_source(self, per_box):
        """A per-box value array expanded to the sources (host array)."""
        return np.asarray(per_box)[self._box_of_source]


def auto_smooth_quad_order(q, split_order, rho_real, rho_imag, config=None):
    """The smooth quadrature order of ``"auto"`` mode (2649-2741):
    max(min, q + (m - 1) * per_order + boosts), boosts from |Im k| h above
    the hard threshold (4) and |Re k| h above 3 (scale 0.5), at most
    ``smooth_quad_order_max``."""
    cfg = dict(config or {})
    min_smooth = int(cfg.get("smooth_quad_order_min", q))
    per = int(cfg.get("smooth_quad_order_per_order", 1))
    per_hard = int(cfg.get("smooth_quad_order_per_order_hard", 1))
    if per < 0 or per_hard < 0:
        raise ValueError("smooth quadrature per-order increments must be non-negative")
    hard_im = float(cfg.get("smooth_quad_order_hard_rho_imag", 4.0))
    hard_re = float(cfg.get("smooth_quad_order_hard_rho_real", 3.0))
    if hard_re <= 0.0:
        raise ValueError("smooth_quad_order_hard_rho_real must be > 0")
    add = per_hard if (rho_imag >= hard_im or rho_real >= hard_re) else per
    b_start = float(cfg.get("smooth_quad_order_rho_boost_start", hard_im))
    b_scale = float(cfg.get("smooth_quad_order_rho_boost_scale", 1.0))
    if b_scale < 0.0:
        raise ValueError("smooth_quad_order_rho_boost_scale must be >= 0")
    boost = int(math.ceil(max(0.0, rho_imag - b_start) * b_scale))
    if cfg.get("smooth_quad_order_rho_boost_cap") is not None:
        boost = min(boost, int(cfg["smooth_quad_order_rho_boost_cap"]))
    r_start = float(cfg.get("smooth_quad_order_real_boost_start", hard_re))
    r_scale = float(cfg.get("smooth_quad_order_real_boost_scale", 0.5))
    if r_scale < 0.0:
        raise ValueError("smooth_quad_order_real_boost_scale must be >= 0")
    rboost = int(math.ceil(max(0.0, rho_real - r_start) * r_scale))
    if cfg.get("smooth_quad_order_real_boost_cap") is not None:
        rboost = min(rboost, int(cfg["smooth_quad_order_real_boost_cap"]))
    order = max(min_smooth, q + max(0, int(split_order) - 1) * add + boost + rboost)
    if cfg.get("smooth_quad_order_max") is not None:
        order = min(order, int(cfg["smooth_quad_order_max"]))
    return order


def split_rho(tree, kernels):
    """(|Re k| h, |Im k| h) for the coarsest active box extent h (4228-4250)."""
    counts = np.asarray(tree.box_target_counts_nonchild)
    active = np.flatnonzero(counts > 0)
    if active.size == 0:
        return 0.0, 0.0
    h = float(tree.root_extent) * 0.5 ** int(np.asarray(tree.box_levels)[active].min())
    kr = max((abs(complex(getattr(kn.get_base_kernel(), "k", 0.0)).real) for kn in kernels),
             default=0.0)
    ki = max((abs(complex(getattr(kn.get_base_kernel(), "k", 0.0)).imag) for kn in kernels),
             default=0.0)
    return kr * h, ki * h


class HelmholtzSplitCorrection:
    """``corr = sum_{list 1} (G_k - G_0)(x_t - y_s) * strength_s`` on the GPU.

    ``k`` is the Helmholtz wave number (complex allowed; Yukawa ``lambda``
    is ``k = 1j * lambda``).  Points, box starts and counts come from the
    tree (sources = targets, tree order).  ``deriv_kind`` /
    ``deriv_axis``: an axis target (``_capi.SPLIT_DERIV_TARGET``) or source
    (``SPLIT_DERIV_SOURCE``) derivative wrapper of the output kernel
    (``expansion_wrangler_fpnd.py:3884-3957``): the pair value becomes
    ``+-(x_a - y_a) R'(r)/r`` and self pairs give 0 (5293-5300).
    ``split_order`` m > 1: the remainder without the terms the split
    tabulates (``split_terms``), also under a derivative wrapper.
    ``legacy_subtraction``: ``G_k - G_0`` per pair from the closed forms
    (Hankel / exponential) instead of the series, the reference's
    ``helmholtz_split_order1_legacy_subtraction`` (3806-3822; order 1),
    self pairs taking the r = 0 limit (3824-3834; 0 under a derivative
    wrapper)."""

    def __init__(self, tree, k, split_order=1, series_terms=None,
                 deriv_kind=_capi.SPLIT_DERIV_NONE, deriv_axis=0, legacy_subtraction=False):
        if int(split_order) < 1:
            raise ValueError("helmholtz_split_order must be >= 1")
        self.legacy = bool(legacy_subtraction)
        if self.legacy and int(split_order) != 1:
            raise NotImplementedError("the legacy subtraction is the split of order 1")
        if self.legacy and complex(k).imag < 0:
            raise ValueError("the legacy subtraction needs Im k >= 0")
        self.dim = int(tree.dim)
        self.k = complex(k)
        self.split_order = int(split_order)
        self.deriv_kind = int(deriv_kind)
        self.deriv_axis = int(deriv_axis)
        if self.deriv_kind not in (_capi.SPLIT_DERIV_NONE, _capi.SPLIT_DERIV_TARGET,
                                   _capi.SPLIT_DERIV_SOURCE):
            raise ValueError("bad deriv_kind")
        if self.deriv_kind and not 0 <= self.deriv_axis < self.dim:
            raise ValueError("deriv_axis out of range")
        r_max = max_neighbor_distance(tree)
        self.nmax = int(series_terms) if series_terms is not None else series_nmax(
            self.dim, self.k, r_max)
        # the series cancels badly at large |k| r: say so instead of
        # returning a silently inaccurate remainder
        self.roundoff = series_roundoff(self.dim, self.k, r_max)
        if self.roundoff > 1e-10 and not self.legacy:
            logger.warning(
                "Helmholtz split remainder: |k| r_max = %.3g; the series loses digits to "
                "cancellation (estimated relative error %.1e of the remainder). Refine the "
                "tree or use per-level Helmholtz tables without the split.",
                abs(self.k) * r_max, self.roundoff)
        self.n_points = int(np.asarray(tree.points).shape[1])
        dev = "cuda"
        self._points = torch.as_tensor(np.ascontiguousarray(tree.points, np.float64), device=dev)
        counts = np.ascontiguousarray(tree.box_target_counts_nonchild, np.int32)
        bstarts = np.ascontiguousarray(tree.box_target_starts, np.int32)
        if counts.size and (counts.min() < 0 or bstarts.min() < 0
                            or int((bstarts.astype(np.int64) + counts).max()) > self.n_points):
            raise ValueError("box starts/counts do not fit the tree's points")
        self.n_boxes = int(counts.size)
        self._starts = torch.as_tensor(bstarts, device=dev)
        self._counts = torch.as_tensor(counts, device=dev)
        self.max_box_points = int(max(1, counts.max(initial=1)))

    def __call__(self, strengths, target_boxes, starts, lists, out=None, stream=None,
                 coefficients=None, sources=None):
        """``strengths[R, n_points]`` (or ``[n_points]``), complex; returns the
        correction with the same shape (targets of unlisted boxes zero, or
        left as they are in ``out``).  ``coefficients``: a radial kernel's
        series coefficients (``vl1_split_args.coefficients`` layout, N =
        ``self.nmax``) evaluated instead of the remainder.  ``sources``: a
        ``SmoothSources`` set; ``strengths`` then index its sources
        (``[R, sources.n_src]``) and the result still lives on the tree's
        points."""
        st = strengths if isinstance(strengths, torch.Tensor) else torch.as_tensor(
            np.asarray(strengths))
        one = st.dim() == 1
        st = (st.unsqueeze(0) if one else st).to(device="cuda", dtype=torch.complex128)
        st = st.contiguous()
        n_src = self.n_points if sources is None else int(sources.n_src)
        if sources is not None and sources.dim != self.dim:
            raise ValueError("source set of another dimension")
        if st.shape[1] != n_src:
            raise ValueError("strengths have the wrong number of points")
        shape = (int(st.shape[0]), self.n_points)
        if out is None:
            out = torch.zeros(shape, dtype=torch.complex128, device="cuda")
        else:
            out = out.unsqueeze(0) if out.dim() == 1 else out
            if tuple(out.shape) != shape or out.dtype != torch.complex128 or not out.is_contiguous():
                raise ValueError("out must be a contiguous complex128 tensor [R, n_points]")
        tb_h = np.ascontiguousarray(_host(target_boxes), np.int32)
        ns_h = np.ascontiguousarray(_host(starts), np.int32)
        nl_h = np.ascontiguousarray(_host(lists), np.int32)
        if ns_h.size != tb_h.size + 1:
            raise ValueError("neighbor_source_boxes_starts must have n_target_boxes + 1 entries")
        if ns_h[0] != 0 or np.any(np.diff(ns_h) < 0) or ns_h[-1] != nl_h.size:
            raise ValueError("neighbor_source_boxes_starts is not a CSR offset array")
        for arr in (tb_h, nl_h):
            if arr.size and (arr.min() < 0 or arr.max() >= self.n_boxes):
                raise ValueError("box index out of range")
        tb = torch.as_tensor(tb_h, device="cuda")
        ns = torch.as_tensor(ns_h, device="cuda")
        nl = torch.as_tensor(nl_h, device="cuda")
        sr = torch.view_as_real(st)
        yr = torch.view_as_real(out)
        args = _capi.SplitArgs(
            dim=self.dim, split_order=self.split_order, series_nmax=self.nmax,
            n_tgt_boxes=int(tb.numel()), n_rhs=int(st.shape[0]),
            max_box_points=self.max_box_points, k_re=self.k.real, k_im=self.k.imag,
            n_points=self.n_points,
            target_boxes=tb.data_ptr(), nbr_starts=ns.data_ptr(),
            nbr_lists=nl.data_ptr() if nl.numel() else ns.data_ptr(),
            box_starts=self._starts.data_ptr(), box_counts=self._counts.data_ptr(),
            points=self._points.data_ptr(), strengths=sr.data_ptr(),
            ld_strengths=2 * n_src, result=yr.data_ptr(), ld_result=2 * self.n_points,
            deriv_kind=self.deriv_kind, deriv_axis=self.deriv_axis,
            legacy_subtraction=int(self.legacy),
        )
        if self.legacy and coefficients is not None:
            raise NotImplementedError(
                "the legacy subtraction sums the closed forms of G_k - G_0 only")
        if sources is not None:
            args.src_points = sources.points.data_ptr()
            args.src_box_starts = sources.box_starts.data_ptr()
            args.src_box_counts = sources.box_counts.data_ptr()
            args.n_src_points = n_src
            args.max_src_box_points = int(sources.m) ** self.dim
        if coefficients is not None:
            cf = np.ascontiguousarray(coefficients, np.float64)
            if self.deriv_kind == _capi.SPLIT_DERIV_NONE:
                want = 2 + 4 * self.nmax if self.dim == 2 else 2 * self.nmax
            else:  # the R'(r)/r layout
                want = 4 * self.nmax if self.dim == 2 else 2 * (self.nmax - 1)
            if cf.shape != (want,):
                raise ValueError(f"coefficients must hold {want} values")
            args.coefficients = cf.ctypes.data
        _capi.check(_capi.lib().vl1_split_correction(args, _capi.stream_handle(stream)),
                    "vl1_split_correction")
        return out[0] if one else out


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)
