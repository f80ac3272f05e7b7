"""Drop-in list-1 part of the FPND expansion wrangler.

``NearFieldWrangler.eval_direct(target_boxes, neighbor_source_boxes_starts,
neighbor_source_boxes_lists, mode_coefs)`` has the signature and return
shape of ``ExpansionWranglerInterface.eval_direct``
(volumential/expansion_wrangler_interface.py:115-128) as implemented by
``FPNDSumpyExpansionWrangler.eval_direct`` (expansion_wrangler_fpnd.py:
3222-3489): it returns ``(pot, timing_future)`` with ``pot`` an object array
holding one potential per output kernel, each overwritten on the targets of
the listed boxes.  Table handling (single table / per-level list / dict by
kernel repr, ``table_starting_level``, coverage and extent checks) follows
2368-2472 and 3254-3273; layout validation follows 408-463.

``eval_direct_batched`` evaluates many densities ``mode_coefs[R, n_points]``
in one call (the per-RHS loop of ``volume_fmm.py:983-1031`` collapsed), and
``eval_direct_plan`` runs the batched plan forms (dense FP64 or Ozaki split).

Helmholtz split (``helmholtz_split``, 2253-2600 and 3491-3836): a Helmholtz
output kernel backed by Laplace tables evaluates its table part with the
Laplace scaling (real and imaginary parts of the density as two real
densities), and ``eval_direct_helmholtz_split_correction`` adds the online
remainder ``G_k - G_0`` over list 1 (``split.HelmholtzSplitCorrection``);
at split order m > 1 also the tabulated terms (``split.split_terms``, each a
plan over its term table).
"""

from __future__ import annotations

import logging
from collections import OrderedDict

import numpy as np
import torch

from . import _capi, gallery
from .kernels import HelmholtzKernel, LaplaceKernel
from .list1 import NearFieldFromCSR
from .payload import prepare_table_data_and_entry_map

logger = logging.getLogger(__name__)


class TimingFuture:
    def __init__(self, seconds=None):
        self.seconds = seconds

    def result(self):
        return {"wall_elapsed": self.seconds, "process_elapsed": None}

    def done(self):
        return True


def _validate_layout(tree, target_boxes, source_boxes, n_q_points):
    """``expansion_wrangler_fpnd.py:427-463``."""
    counts = np.asarray(tree.box_target_counts_nonchild)
    tc = counts[np.asarray(target_boxes, dtype=np.int64)]
    sc = counts[np.asarray(source_boxes, dtype=np.int64)]
    if np.all(tc == n_q_points) and np.all(sc == n_q_points):
        return
    bad_t = np.unique(tc[tc != n_q_points])
    bad_s = np.unique(sc[sc != n_q_points])
    raise ValueError(
        "table-based near-field evaluation requires exactly "
        f"{n_q_points} quadrature points per active source/target box; "
        f"found target counts {bad_t.tolist()} and source counts {bad_s.tolist()}."
    )


def _validate_local_ids(tree, boxes, n):
    """Each listed box's node ids must be a permutation of 0..n-1."""
    starts = np.asarray(tree.box_target_starts, dtype=np.int64)[boxes]
    ids = np.asarray(tree.local_ids)
    blk = ids[starts[:, None] + np.arange(n)[None, :]]
    if not np.array_equal(np.sort(blk, axis=1), np.broadcast_to(np.arange(n), blk.shape)):
        raise ValueError("box-local node ids must be a permutation of 0..q^d-1 in every box")


class NearFieldWrangler:
    def __init__(
        self,
        tree,
        near_field_table,
        target_kernels=None,
        dtype=np.float64,
        list1_extra_kwargs=None,
        helmholtz_split=None,
        direct_method="plan",
        helmholtz_split_order=1,
        helmholtz_split_auto_config=None,
        helmholtz_split_term_tables=None,
        helmholtz_split_term_build_config=None,
        helmholtz_split_smooth_quad_order=None,
        helmholtz_split_order1_legacy_subtraction=False,
        potential_kind=1,
    ):
        """``direct_method``: how ``eval_direct`` evaluates list 1 --
        ``"plan"`` (default) through a plan cached per list-1 CSR, FP64
        MFMA form (the same sums as the reference kernel to ~1e-15, at a
        fraction of the per-call cost); ``"plain"`` through the plain FP64
        kernel with the reference's per-call semantics (``NearFieldFromCSR``).

        ``helmholtz_split``: ``None`` enables the split for Helmholtz output
        kernels whose tables are Laplace tables (``expansion_wrangler_fpnd.py:
        2482-2520``), ``True`` requires it, ``False`` disables it.

        ``helmholtz_split_order`` (2289-2310, 2520-2555): an integer m >= 1 or
        ``"auto"`` (``helmholtz_split_auto_config``, ``split.choose_auto_split_order``).
        An order m > 1 tabulates the non-smooth terms of the remainder (2D
        r^2n log r, 3D r^(2j-1), n, j < m) and leaves the rest to the online
        sum.  ``helmholtz_split_term_tables`` maps ``("power", p)`` /
        ``("power_log", p)`` to tables (one per Laplace table level); missing
        ones are built on the GPU with ``helmholtz_split_term_build_config``
        (a ``DuffyBuildConfig``, default rule).  Several output kernels run at
        order 1, as in the reference (2558-2568).

        ``helmholtz_split_smooth_quad_order`` (2303-2306, 2649-2741): the
        tensor Gauss order of the online remainder's sources; above q the
        correction runs on those finer per-box nodes with the density
        interpolated to them (given ``src_func``, 3562-3597).  Default: q, or
        in ``"auto"`` mode the reference's rule (``split.auto_smooth_quad_order``).

        ``helmholtz_split_auto_config["power_log_single_table_beta_mode"]``
        (2297-2301, 3709-3786): how the log(h/H) part of a 2D r^2n log r term
        from one table is summed -- ``"p2p"`` (default: a point-to-point sum
        of r^2n over the strengths times log(h_s/H)) or ``"table"`` (the r^2n
        term table applied to the density times log(h_s/H)).

        ``helmholtz_split_order1_legacy_subtraction`` (2313-2316, 3806-3822):
        at split order 1 the online correction sums ``G_k - G_0`` (or its
        axis derivative) from the closed forms (2D Hankel, 3D exponential)
        instead of the series; self pairs take the r = 0 limit; a smooth
        quadrature order whose nodes meet the box nodes is refused
        (5125-5138).  The fused split plan (series quadrature in its blocks)
        is not taken.

        ``potential_kind`` (2257, 2407, 3329): 2 for the inverse-type
        potentials, whose list 1 adds the exterior mode normalisers of the
        tables; evaluated by the plain route (``NearFieldFromCSR``)."""
        if potential_kind not in (1, 2):
            raise ValueError("Unsupported potential kind %d" % potential_kind)
        self.potential_kind = int(potential_kind)
        if not getattr(tree, "sources_are_targets", False):
            raise ValueError(
                "table-based near-field evaluation requires sources and targets "
                "to coincide (tree.sources_are_targets=True)"
            )
        if direct_method not in ("plan", "plain"):
            raise ValueError("direct_method must be 'plan' or 'plain'")
        self.direct_method = direct_method
        self.tree = tree
        self.dtype = np.dtype(dtype)
        if isinstance(near_field_table, dict):
            tables = {k: self._as_list(v) for k, v in near_field_table.items()}
        else:
            tables = None
            lst = self._as_list(near_field_table)
        if target_kernels is None:
            if tables is not None:
                raise ValueError("target_kernels is required with a dict of tables")
            target_kernels = [lst[0].integral_knl]
        self.target_kernels = list(target_kernels)
        if tables is None:
            tables = {repr(k): lst for k in self.target_kernels}
        self.near_field_table = tables
        self.list1_extra_kwargs = dict(list1_extra_kwargs or {})
        for kname, lst in tables.items():
            t0 = lst[0]
            if t0.n_q_points != tree.n_q_points:
                raise ValueError("table order does not match the tree's nodes per box")
            lvl0 = self.table_starting_level(lst)
            for lev, t in enumerate(lst):
                if not np.isclose(t.source_box_extent * 2**lev, t0.source_box_extent):
                    raise RuntimeError(
                        f"table level {lev} extent {t.source_box_extent} is inconsistent with "
                        f"level-0 extent {t0.source_box_extent}"
                    )
            if len(lst) > 1 and tree.nlevels > len(lst) + lvl0:
                raise RuntimeError(
                    f"Insufficient list of tables: the tree has {tree.nlevels} levels but "
                    f"only {len(lst)} tables starting at level {lvl0} were given."
                )
        self._evaluators = {}
        # plans by (kernel, CSR, fused): least recently used first; each holds
        # its expanded blocks on the device, so the cache is bounded
        self._plans = OrderedDict()
        self.max_cached_plans = 8
        self._corrections = {}
        self._term_tables = {}
        self._term_plans = OrderedDict()
        self._beta = {}
        self.helmholtz_split_term_build_config = helmholtz_split_term_build_config
        for key, tabs in (helmholtz_split_term_tables or {}).items():
            kind, p = key
            if kind not in ("power", "power_log"):
                raise ValueError(f"unknown split term kind {kind!r}")
            self._term_tables[(kind, int(p))] = self._as_list(tabs)
        self.helmholtz_split_order = 1
        self.helmholtz_split_beta_mode = str(
            dict(helmholtz_split_auto_config or {}).get("power_log_single_table_beta_mode",
                                                        "p2p")).strip().lower()
        if self.helmholtz_split_beta_mode not in ("p2p", "table"):
            raise ValueError("power_log_single_table_beta_mode must be 'table' or 'p2p'")
        self.helmholtz_split = False
        self.helmholtz_split_order1_legacy_subtraction = bool(
            helmholtz_split_order1_legacy_subtraction)
        self.helmholtz_split_smooth_quad_order = None
        self._smooth_sources = {}
        if helmholtz_split is None or helmholtz_split:
            ok, why = self._split_support_status()
            if ok:
                self.helmholtz_split = True
                self.helmholtz_split_order = self._resolve_split_order(
                    helmholtz_split_order, helmholtz_split_auto_config)
                self.helmholtz_split_smooth_quad_order = self._resolve_smooth_order(
                    helmholtz_split_smooth_quad_order, helmholtz_split_order,
                    helmholtz_split_auto_config)
            elif helmholtz_split:
                raise RuntimeError(f"helmholtz_split is not supported here: {why}")
        if self.helmholtz_split and self._legacy():
            ms = self.helmholtz_split_smooth_quad_order or 0
            if ms > int(self.tree.quad_order):
                # the closed forms are singular where a smooth node meets a
                # box node (5125-5138)
                from .quadrature import gauss_legendre_01

                xq, _ = gauss_legendre_01(int(self.tree.quad_order))
                xm, _ = gauss_legendre_01(ms)
                if np.any(np.abs(xm[:, None] - xq[None, :]) < 1.0e-15):
                    raise ValueError(
                        "helmholtz split smooth quadrature order shares nodes with source "
                        "quadrature; choose a non-overlapping order to avoid singular "
                        "Helmholtz/Laplace evaluations")
        self._plain_only = set()  # kernels the plan route cannot take

    def _resolve_split_order(self, order, auto_config):
        """``expansion_wrangler_fpnd.py:2520-2568``; in "auto" mode an
        imaginary rho above ``rho_imag_split_max`` (default 8) is reported,
        and the split stays on (2550-2568)."""
        from .split import choose_auto_split_order, split_rho

        cfg = dict(auto_config or {})
        auto = bool(cfg.get("enabled", False)) or (
            isinstance(order, str) and order.strip().lower() == "auto")
        if auto:
            order = choose_auto_split_order(self.tree, self.target_kernels, cfg)
            logger.info("Auto-selected helmholtz_split_order=%d", order)
        order = int(order)
        if order < 1:
            raise ValueError("helmholtz_split_order must be >= 1")
        if order > 1 and len(self.target_kernels) > 1:
            logger.warning("split order %d requested with %d target kernels; clamping to order "
                           "1 for multi-output split mode", order, len(self.target_kernels))
            order = 1
        if auto:
            rho_im = split_rho(self.tree, self.target_kernels)[1]
            max_im = float(cfg.get("rho_imag_split_max", 8.0))
            if rho_im > max_im:
                if bool(cfg.get("disable_split_if_outside_coverage", False)):
                    logger.warning(
                        "Imaginary rho %.3g exceeds configured split coverage %.3g. Keeping "
                        "split mode enabled because direct fallback requires matching direct "
                        "near-field tables.", rho_im, max_im)
                else:
                    logger.warning(
                        "Imaginary rho %.3g exceeds configured split coverage %.3g; continuing "
                        "with split at configured max order", rho_im, max_im)
        return order

    def _resolve_smooth_order(self, smooth_order, order_arg, auto_config):
        """``expansion_wrangler_fpnd.py:2649-2747``."""
        from .split import auto_smooth_quad_order, split_rho

        cfg = dict(auto_config or {})
        auto = bool(cfg.get("enabled", False)) or (
            isinstance(order_arg, str) and order_arg.strip().lower() == "auto")
        if smooth_order is None:
            if auto:
                rho_re, rho_im = split_rho(self.tree, self.target_kernels)
                smooth_order = auto_smooth_quad_order(
                    int(self.tree.quad_order), self.helmholtz_split_order, rho_re, rho_im, cfg)
            else:
                smooth_order = int(self.tree.quad_order)
        smooth_order = int(smooth_order)
        if smooth_order < 1:
            raise ValueError("helmholtz_split_smooth_quad_order must be >= 1")
        return smooth_order

    def smooth_sources(self, m):
        """The smooth-quadrature source set of order ``m`` (cached)."""
        from .split import SmoothSources

        if m not in self._smooth_sources:
            self._smooth_sources[m] = SmoothSources(self.tree, m)
        return self._smooth_sources[m]

    def _split_support_status(self):
        """``_split_target_kernel_support_status`` /
        ``_split_base_table_support_status`` (4009-4073, 3964-4007) for the
        kernels of this library: Helmholtz output kernels, bare or under one
        axis target / source derivative wrapper, whose tables all hold the
        same wrapper around ``LaplaceKernel(dim)``."""
        for knl in self.target_kernels:
            try:
                base, _, _ = split_wrapper_chain(knl)
            except NotImplementedError as exc:
                return False, str(exc)
            if not isinstance(base, HelmholtzKernel):
                return False, f"{knl!r} is not a Helmholtz kernel"
            want = repr(_rewrap(knl, LaplaceKernel(knl.dim)))
            for lev, t in enumerate(self.near_field_table.get(repr(knl), [])):
                if repr(t.integral_knl) != want:
                    return False, (f"{knl!r} level {lev} kernel mismatch: expected {want}, "
                                   f"got {t.integral_knl!r}")
        return True, ""

    def _policy_kernel(self, out_kernel):
        """The kernel whose scaling rule the table part follows: the tables'
        Laplace kernel under the split (4126-4153, 4801-4833), else the
        output kernel."""
        if self.helmholtz_split:
            return self.near_field_table[repr(out_kernel)][0].integral_knl
        return out_kernel

    @staticmethod
    def _as_list(t):
        return list(t) if isinstance(t, (list, tuple)) else [t]

    def table_starting_level(self, tables):
        return int(round(np.log2(self.tree.root_extent / tables[0].source_box_extent)))

    @staticmethod
    def _dtype(kernel):
        """complex128 for complex kernels (``expansion_wrangler_fpnd.py:3283-3297``)."""
        return torch.complex128 if getattr(kernel, "is_complex_valued", False) else torch.float64

    def _finish(self, pot):
        """Round results to the wrangler's single precision when it has one:
        ``expansion_wrangler_fpnd.py:3283-3297`` evaluates in f32 / c64 for a
        float32 wrangler; this library evaluates in FP64 and rounds the result
        (it agrees with an f32 evaluation to f32 rounding, and is closer to
        the exact sum)."""
        if self.dtype not in (np.dtype(np.float32), np.dtype(np.complex64)):
            return pot
        for i in range(len(pot)):
            pot[i] = pot[i].to(torch.complex64 if pot[i].is_complex() else torch.float32)
        return pot

    def output_zeros(self, n_rhs=None):
        shape = (self.tree.n_points,) if n_rhs is None else (n_rhs, self.tree.n_points)
        pot = np.empty(len(self.target_kernels), dtype=object)
        for i, knl in enumerate(self.target_kernels):
            pot[i] = torch.zeros(shape, dtype=self._dtype(knl), device="cuda")
        return pot

    # -------------------------------------------------------------- eval
    def _eval_single(self, out_pot, out_kernel, target_boxes, starts, lists, mode_coefs):
        if self.helmholtz_split:
            # real Laplace tables: real and imaginary parts as two real densities
            mc = mode_coefs if isinstance(mode_coefs, torch.Tensor) else torch.as_tensor(
                np.asarray(mode_coefs))
            mc = mc.to(device="cuda", dtype=torch.complex128)
            one = mc.dim() == 1
            m2 = mc.unsqueeze(0) if one else mc
            r = int(m2.shape[0])
            stacked = torch.cat([m2.real, m2.imag], 0).contiguous()
            tmp = torch.zeros_like(stacked)
            self._eval_real(tmp, out_kernel, target_boxes, starts, lists, stacked)
            res = torch.complex(tmp[:r], tmp[r:])
            out_pot.copy_(res[0] if one else res)
            return out_pot
        return self._eval_real(out_pot, out_kernel, target_boxes, starts, lists, mode_coefs)

    def _eval_real(self, out_pot, out_kernel, target_boxes, starts, lists, mode_coefs):
        kname = repr(out_kernel)
        tables = self.near_field_table[kname]
        t0 = tables[0]
        tsl = self.table_starting_level(tables)
        if len(tables) > 1:
            min_lev = int(np.min(np.asarray(self.tree.box_levels)[np.asarray(target_boxes)]))
            largest = self.tree.root_extent * 0.5**min_lev
            if not t0.source_box_extent >= largest - 1e-15:
                raise RuntimeError(
                    "Insufficient list of tables: the coarsest level mesh cells at level "
                    f"{min_lev} are not covered."
                )
        if (self.direct_method == "plan" and kname not in self._plain_only
                and self.potential_kind == 1):
            key = _csr_key(out_kernel, target_boxes, starts, lists) + (False,)
            if key not in self._plans:  # host layout checks, once per CSR
                _validate_layout(self.tree, np.asarray(_host(target_boxes), dtype=np.int64),
                                 np.asarray(_host(lists), dtype=np.int64), t0.n_q_points)
                try:
                    self.get_plan(out_kernel, target_boxes, starts, lists)
                except NotImplementedError:
                    # box sizes / the complex 2D log term the tile kernels do not
                    # take: the plain kernel evaluates this kernel from now on
                    self._plain_only.add(kname)
            if kname not in self._plain_only:
                return self._eval_plan(out_pot, out_kernel, target_boxes, starts, lists,
                                       mode_coefs)
        data, nmlz, ext, _, _, info = prepare_table_data_and_entry_map(tables)
        shapes = {
            "n_tables": len(tables),
            "n_q_points": t0.n_q_points,
            "n_table_entries": data.shape[1],
            "n_cases": t0.n_cases,
            "quad_order": t0.quad_order,
            "reconstruction_kind": info["kind"],
            "n_arithmetic_case_orbits": info["n_case_orbits"],
        }
        key = (kname, len(tables))
        if key not in self._evaluators:
            self._evaluators[key] = NearFieldFromCSR(self._policy_kernel(out_kernel), shapes,
                                                     potential_kind=self.potential_kind,
                                                     **self.list1_extra_kwargs)
        nf = self._evaluators[key]
        tb = np.asarray(_host(target_boxes), dtype=np.int64)
        ls = np.asarray(_host(lists), dtype=np.int64)
        _validate_layout(self.tree, tb, ls, t0.n_q_points)
        _validate_local_ids(self.tree, np.unique(np.concatenate([tb, ls])), t0.n_q_points)
        tree = self.tree
        res, _ = nf(
            None,
            result=out_pot,
            box_centers=tree.box_centers,
            box_levels=tree.box_levels,
            box_source_counts_nonchild=tree.box_target_counts_nonchild,
            box_source_starts=tree.box_target_starts,
            box_target_counts_nonchild=tree.box_target_counts_nonchild,
            box_target_starts=tree.box_target_starts,
            case_indices=t0.case_indices,
            encoding_base=gallery.ENCODING_BASE,
            encoding_shift=gallery.ENCODING_SHIFT,
            mode_nmlz_combined=nmlz,
            exterior_mode_nmlz_combined=ext,
            neighbor_source_boxes_starts=starts,
            root_extent=tree.root_extent,
            neighbor_source_boxes_lists=lists,
            mode_coefs=mode_coefs,
            source_mode_ids=tree.local_ids,
            table_data_combined=data,
            target_boxes=target_boxes,
            target_point_ids=tree.local_ids,
            table_root_extent=t0.source_box_extent,
            table_starting_level=tsl,
            arithmetic_case_orbit_ranks=info["case_orbit_ranks"],
            arithmetic_case_axis_perm=info["case_axis_perm"],
            arithmetic_case_axis_sign=info["case_axis_sign"],
            arithmetic_case_axis_group=info["case_axis_group"],
            arithmetic_case_value_offsets=info["case_value_offsets"],
            arithmetic_axis_sign_power=info["axis_sign_power"],
        )
        return res

    def _eval_plan(self, out_pot, out_kernel, target_boxes, starts, lists, mode_coefs):
        """FP64 plan evaluation into ``out_pot`` (targets of listed boxes
        overwritten), in RHS chunks the tile kernels take (<= 64, or whole
        multiples of 64)."""
        mc = mode_coefs if isinstance(mode_coefs, torch.Tensor) else torch.as_tensor(
            np.asarray(mode_coefs))
        mc = mc.to(device="cuda", dtype=self._dtype(self._policy_kernel(out_kernel)))
        one = mc.dim() == 1
        mc = (mc.unsqueeze(0) if one else mc).contiguous()
        out = out_pot.unsqueeze(0) if one else out_pot
        plan = self.get_plan(out_kernel, target_boxes, starts, lists)
        r = int(mc.shape[0])
        bulk = r - r % 64 if r > 64 else r
        plan.eval(mc[:bulk], out=out[:bulk], form="f64")
        if bulk < r:
            plan.eval(mc[bulk:], out=out[bulk:], form="f64")
        return out_pot

    def eval_direct(self, target_boxes, neighbor_source_boxes_starts, neighbor_source_boxes_lists,
                    mode_coefs):
        pot = self.output_zeros()
        for i, knl in enumerate(self.target_kernels):
            pot[i] = self._eval_single(
                pot[i], knl, target_boxes, neighbor_source_boxes_starts,
                neighbor_source_boxes_lists, mode_coefs,
            )
        return self._finish(pot), TimingFuture()

    def eval_direct_batched(self, target_boxes, neighbor_source_boxes_starts,
                            neighbor_source_boxes_lists, mode_coefs):
        """``mode_coefs[R, n_points]`` -> ``pot[i][R, n_points]``."""
        mc = mode_coefs if isinstance(mode_coefs, torch.Tensor) else torch.as_tensor(
            np.asarray(mode_coefs)
        )
        if mc.dim() != 2:
            raise ValueError("mode_coefs must have shape (n_rhs, n_points)")
        pot = self.output_zeros(int(mc.shape[0]))
        for i, knl in enumerate(self.target_kernels):
            pot[i] = self._eval_single(
                pot[i], knl, target_boxes, neighbor_source_boxes_starts,
                neighbor_source_boxes_lists, mc,
            )
        return self._finish(pot), TimingFuture()


    # ------------------------------------------------------------ plans
    def get_plan(self, out_kernel, target_boxes, neighbor_source_boxes_starts,
                 neighbor_source_boxes_lists, fused_split=False):
        """Plan for one output kernel and one list-1 CSR (cached).
        ``fused_split``: the Helmholtz split plan (table part + correction in
        one contraction, ``vl1_plan_create_helmholtz_split``)."""
        from .payload import DevicePayload
        from .plan import List1Plan

        if self.potential_kind != 1:
            raise NotImplementedError(
                "plans evaluate potential kind 1; potential_kind=2 runs through eval_direct "
                "(the plain route)")
        key = _csr_key(out_kernel, target_boxes, neighbor_source_boxes_starts,
                       neighbor_source_boxes_lists) + (bool(fused_split),)
        if key in self._plans:
            self._plans.move_to_end(key)
            return self._plans[key]
        tables = self.near_field_table[repr(out_kernel)]
        t0 = tables[0]
        data, nmlz, _, _, _, info = prepare_table_data_and_entry_map(tables)
        payload = DevicePayload(t0.dim, t0.quad_order, data, nmlz, t0.case_indices,
                                t0.arithmetic_metadata())
        nf = NearFieldFromCSR(
            self._policy_kernel(out_kernel),
            {"n_tables": len(tables), "n_q_points": t0.n_q_points,
             "n_table_entries": data.shape[1], "n_cases": t0.n_cases,
             "quad_order": t0.quad_order, "reconstruction_kind": info["kind"]},
            **self.list1_extra_kwargs,
        )
        pol = nf.get_kernel_scaling_policy()
        # the reference checks the drift on every call; a plan checks its CSR once
        from .list1 import warn_case_encoding_drift

        warn_case_encoding_drift(
            self.tree.dim, self.tree.box_centers, self.tree.box_levels, self.tree.root_extent,
            np.frombuffer(key[1], dtype=np.int32), np.frombuffer(key[2], dtype=np.int32),
            np.frombuffer(key[3], dtype=np.int32), gallery.ENCODING_SHIFT,
            nf.extra_kwargs.get("case_encoding_warn_tol"))
        lsc, ldisp = nf.custom_level_arrays(self.tree.root_extent, t0.source_box_extent)
        split_desc = None
        if fused_split:
            corr = self._correction(out_kernel)
            split_desc = {"k": corr.k, "series_nmax": corr.nmax,
                          "deriv_kind": corr.deriv_kind, "deriv_axis": corr.deriv_axis,
                          "split_order": self.helmholtz_split_order}
            if self.helmholtz_split_order > 1:
                from .split import split_terms

                base, _, _ = split_wrapper_chain(out_kernel)
                terms = []
                for kind, p, coeff in split_terms(self.tree.dim, base.k,
                                                  self.helmholtz_split_order):
                    tabs = self.split_term_tables(out_kernel, (kind, int(p)))
                    tdata = prepare_table_data_and_entry_map(tabs)[0]
                    terms.append((np.asarray(tdata[0]), int(p), kind == "power_log", coeff))
                split_desc["terms"] = terms
            ms = self.helmholtz_split_smooth_quad_order or 0
            if ms > int(self.tree.quad_order):
                from .quadrature import gauss_legendre_01
                from .split import interp_matrix

                xq, _ = gauss_legendre_01(int(self.tree.quad_order))
                xm, wm = gauss_legendre_01(ms)
                split_desc["smooth"] = (xm, wm, interp_matrix(xq, xm))
        plan = List1Plan(
            self.tree, payload,
            scaling_mode=nf._mode_code(pol),
            scaling_power=pol.scaling_power,
            log_displacement=pol.log_displacement,
            table_root_extent=t0.source_box_extent,
            table_starting_level=self.table_starting_level(tables),
            case_encoding_bias=nf.extra_kwargs["case_encoding_bias"],
            target_boxes=np.frombuffer(key[1], dtype=np.int32),
            starts=np.frombuffer(key[2], dtype=np.int32),
            lists=np.frombuffer(key[3], dtype=np.int32),
            helmholtz_split=split_desc,
            level_scaling=lsc, level_disp=ldisp,
        )
        self._plans[key] = plan
        while len(self._plans) > self.max_cached_plans:
            # dropped from the cache, not destroyed: a captured graph may still
            # hold the plan (its device memory is freed with the last reference)
            self._plans.popitem(last=False)
        return plan

    def eval_direct_plan(self, target_boxes, neighbor_source_boxes_starts,
                         neighbor_source_boxes_lists, mode_coefs, form="split", n_slices=None,
                         tol=1e-10):
        """Batched evaluation of ``mode_coefs[R, n_points]`` with a plan form
        ("f64", "split" (BF16) or "split_i8"); returns ``(pot,
        timing_future)`` like eval_direct.

        ``form="split_i8"`` is accuracy-guarded: every evaluation comes with
        its per-RHS error estimate (relative Linf and L2 over the targets);
        with ``n_slices=None`` a density whose estimate exceeds ``tol`` is
        evaluated again with one more slice (5, then 6), and one still above
        ``tol`` at 6 slices with the FP64 form.  The slices used, the
        estimates and the FP64 fallbacks of the last call are in
        ``self.last_split_report``."""
        mc0 = mode_coefs if isinstance(mode_coefs, torch.Tensor) else torch.as_tensor(
            np.asarray(mode_coefs))
        if mc0.dim() == 1:
            mc0 = mc0.unsqueeze(0)
        pot = self.output_zeros(int(mc0.shape[0]))
        self.last_split_report = []
        for i, knl in enumerate(self.target_kernels):
            plan = self.get_plan(knl, target_boxes, neighbor_source_boxes_starts,
                                 neighbor_source_boxes_lists)
            if self.helmholtz_split:  # real Laplace plan on [Re f; Im f]
                mc = mc0.to(device="cuda", dtype=torch.complex128)
                r = int(mc.shape[0])
                stacked = torch.cat([mc.real, mc.imag], 0).contiguous()
                tmp = torch.zeros_like(stacked)
                self._plan_eval(plan, stacked, tmp, form, n_slices, tol)
                pot[i].copy_(torch.complex(tmp[:r], tmp[r:]))
                continue
            mc = mc0.to(device="cuda", dtype=self._dtype(knl)).contiguous()
            self._plan_eval(plan, mc, pot[i], form, n_slices, tol)
        return self._finish(pot), TimingFuture()

    def _plan_eval(self, plan, mc, out, form, n_slices, tol):
        if form != "split_i8":
            plan.eval(mc, out=out, form=form, n_slices=n_slices)
            return
        R = int(mc.shape[0])
        est = torch.zeros((R, 2), dtype=torch.float64, device="cuda")
        S = 5 if n_slices is None else int(n_slices)
        plan.eval(mc, out=out, form="split_i8", n_slices=S, error_est=est)
        e = est.cpu().numpy()
        slices = np.full(R, S, dtype=np.int32)
        bad = np.flatnonzero(e.max(axis=1) > tol)
        while bad.size and n_slices is None and S < 6:
            S += 1
            for c0 in range(0, bad.size, 64):  # plans take <= 64 RHS or multiples of 64
                idx = bad[c0:c0 + 64]
                sub, sub_out = mc[idx].contiguous(), out[idx].contiguous()
                sub_est = torch.zeros((idx.size, 2), dtype=torch.float64, device="cuda")
                plan.eval(sub, out=sub_out, form="split_i8", n_slices=S, error_est=sub_est)
                out[idx] = sub_out
                e[idx] = sub_est.cpu().numpy()
            slices[bad] = S
            bad = bad[e[bad].max(axis=1) > tol]
        fallback = np.zeros(R, dtype=bool)
        if bad.size and n_slices is None:
            for c0 in range(0, bad.size, 64):
                idx = bad[c0:c0 + 64]
                sub_out = out[idx].contiguous()
                plan.eval(mc[idx].contiguous(), out=sub_out, form="f64")
                out[idx] = sub_out
            fallback[bad] = True
        self.last_split_report.append({"slices": slices, "estimate": e, "fallback_f64": fallback})

    # --------------------------------------------------------- split
    def _correction(self, knl):
        key = repr(knl)
        if key not in self._corrections:
            from .split import HelmholtzSplitCorrection

            base, kind, axis = split_wrapper_chain(knl)
            self._corrections[key] = HelmholtzSplitCorrection(
                self.tree, base.k, split_order=self.helmholtz_split_order,
                deriv_kind=kind, deriv_axis=axis, legacy_subtraction=self._legacy())
        return self._corrections[key]

    def _legacy(self):
        """The legacy subtraction applies at split order 1 only (3573-3578)."""
        return self.helmholtz_split_order1_legacy_subtraction and self.helmholtz_split_order == 1

    def eval_direct_helmholtz_split_correction(self, target_boxes, neighbor_source_boxes_starts,
                                               neighbor_source_boxes_lists, src_weights,
                                               src_func=None):
        """The online list-1 remainder ``sum (G_k - G_0)(x_t - y_s) w_s`` per
        output kernel (``expansion_wrangler_fpnd.py:3491-3836``), with
        ``src_weights`` = density x quadrature weight (``[n_points]`` or
        ``[R, n_points]``).  Zeros when the split is off.  This library
        implements the base-quadrature path.  At split order m > 1 the
        remainder leaves out the tabulated terms, and their table
        evaluations on ``src_func`` (the density values, required then, as in
        the reference, 3580-3586) are added."""
        if not self.helmholtz_split:
            return self._finish(self.output_zeros()), TimingFuture()
        q = int(self.tree.quad_order)
        ms = self.helmholtz_split_smooth_quad_order
        smooth = ms is not None and ms > q and src_func is not None
        if self.helmholtz_split_order > 1 and not (
                smooth or (ms == q and src_func is not None)):
            raise RuntimeError("helmholtz_split_order > 1 requires source function samples "
                               "(src_func) and smooth quadrature order >= q")
        w = src_weights if isinstance(src_weights, torch.Tensor) else torch.as_tensor(
            np.asarray(src_weights))
        sources, st = None, w
        if smooth:  # interpolated density on the finer per-box nodes (5188-5275)
            sources = self.smooth_sources(ms)
            f = src_func if isinstance(src_func, torch.Tensor) else torch.as_tensor(
                np.asarray(src_func))
            f = f.to(device="cuda", dtype=torch.complex128)
            one = f.dim() == 1
            st = sources.strengths(f.unsqueeze(0) if one else f)
            if one:
                st = st[0]
        pot = np.empty(len(self.target_kernels), dtype=object)
        for i, knl in enumerate(self.target_kernels):
            pot[i] = self._correction(knl)(st, target_boxes, neighbor_source_boxes_starts,
                                           neighbor_source_boxes_lists, sources=sources)
            if self.helmholtz_split_order > 1:
                pot[i] = pot[i] + self._split_terms_value(
                    knl, target_boxes, neighbor_source_boxes_starts, neighbor_source_boxes_lists,
                    src_func, st, sources)
        return self._finish(pot), TimingFuture()

    # ---------------------------------------------- split terms (order > 1)
    def split_term_tables(self, knl, key):
        """Tables of split term ``key`` = ``("power", p)`` / ``("power_log",
        p)``: the caller's, or built on the GPU at the extents of ``knl``'s
        Laplace tables (``_autobuild_helmholtz_split_term_tables``,
        4700-4776); under ``knl``'s derivative wrapper for a wrapped output
        kernel."""
        from .kernels import RadialPowerKernel, RadialPowerLogKernel

        kind, p = key
        dim = self.tree.dim
        # the term under the output kernel's derivative wrapper (5018-5051)
        term = _rewrap(knl, RadialPowerKernel(dim, p) if kind == "power"
                       else RadialPowerLogKernel(dim, p))
        if split_wrapper_chain(knl)[1] != _capi.SPLIT_DERIV_NONE:
            key = (kind, p, repr(term))  # the caller's tables are the bare terms'
        if key not in self._term_tables:
            from .table import DuffyBuildConfig, NearFieldInteractionTable

            cfg = self.helmholtz_split_term_build_config or DuffyBuildConfig()
            built = []
            for t in self.near_field_table[repr(knl)]:
                tt = NearFieldInteractionTable(t.quad_order, dim, term,
                                               source_box_extent=t.source_box_extent,
                                               source_box_level=t.source_box_level)
                built.append(tt.build_table(cfg))
            self._term_tables[key] = built
        tabs = self._term_tables[key]
        lap = self.near_field_table[repr(knl)]
        if len(tabs) != len(lap) or any(
                not np.isclose(a.source_box_extent, b.source_box_extent) for a, b in zip(tabs, lap)):
            raise RuntimeError(f"split term tables {key} do not match the Laplace table levels")
        return tabs

    def _term_plan(self, knl, key, target_boxes, starts, lists, log_factor=False):
        """A plan over term tables ``key``: one table scaled by
        (h/H)^(d + p) (``_helmholtz_split_term_scaling_code``, 4777-4798),
        or per-level tables unscaled.  ``log_factor``: one table scaled by
        (h/H)^(d + p) log(h/H) -- the "table" form of a 2D log term's
        log(h/H) part (3757-3779), the factor of the source box's level."""
        from .payload import DevicePayload
        from .plan import List1Plan

        ck = (("term", bool(log_factor)) + tuple(key)
              + _csr_key(knl, target_boxes, starts, lists)[1:])
        if ck in self._term_plans:
            self._term_plans.move_to_end(ck)
            return self._term_plans[ck]
        tabs = self.split_term_tables(knl, key)
        t0 = tabs[0]
        data, nmlz, _, _, _, _ = prepare_table_data_and_entry_map(tabs)
        payload = DevicePayload(t0.dim, t0.quad_order, data, nmlz, t0.case_indices,
                                t0.arithmetic_metadata())
        kw = {}
        if len(tabs) == 1:
            # (h/H)^(d + p), one power less under a derivative wrapper
            n_deriv = int(split_wrapper_chain(knl)[1] != _capi.SPLIT_DERIV_NONE)
            e = self.tree.dim + int(key[1]) - n_deriv
            lev = np.arange(_capi.MAX_LEVELS)
            h = float(self.tree.root_extent) * 0.5 ** lev
            scal = (h / t0.source_box_extent) ** e
            if log_factor:
                scal = scal * np.log(h / t0.source_box_extent)
            kw = dict(scaling_mode=_capi.SCALING_CUSTOM, level_scaling=scal,
                      level_disp=np.zeros(_capi.MAX_LEVELS))
        elif log_factor:
            raise RuntimeError("the log(h/H) factor applies to single-table terms only")
        else:
            kw = dict(scaling_mode=_capi.SCALING_PER_LEVEL)
        plan = List1Plan(
            self.tree, payload, table_root_extent=t0.source_box_extent,
            table_starting_level=self.table_starting_level(tabs),
            target_boxes=np.ascontiguousarray(_host(target_boxes), np.int32),
            starts=np.ascontiguousarray(_host(starts), np.int32),
            lists=np.ascontiguousarray(_host(lists), np.int32), **kw)
        self._term_plans[ck] = plan
        while len(self._term_plans) > self.max_cached_plans:
            self._term_plans.popitem(last=False)
        return plan

    def _split_terms_value(self, knl, target_boxes, starts, lists, src_func, strengths,
                           sources=None):
        """sum over the split's terms of coefficient x (term table on
        ``src_func``), with, for a 2D log term from one table, its
        log(h/H) part: the point-to-point sum of r^2n over the strengths
        times log(h_s/H) of each source's box (``beta``, 3689-3800), or in
        the "table" beta mode the r^2n term table on ``src_func`` times
        log(h_s/H) (3757-3779)."""
        from .split import split_terms

        base, _, _ = split_wrapper_chain(knl)
        f = src_func if isinstance(src_func, torch.Tensor) else torch.as_tensor(
            np.asarray(src_func))
        f = f.to(device="cuda", dtype=torch.complex128)
        one = f.dim() == 1
        f2 = (f.unsqueeze(0) if one else f).contiguous()
        r = int(f2.shape[0])
        # real term tables: the real and imaginary parts as 2R real rows
        stacked = torch.cat([f2.real, f2.imag], 0).contiguous()
        rows = 2 * r
        bulk = rows - rows % 64 if rows > 64 else rows  # plans take <= 64 rows or multiples
        st = strengths.to(device="cuda", dtype=torch.complex128)
        st = (st.unsqueeze(0) if st.dim() == 1 else st).contiguous()
        total = torch.zeros_like(f2)
        corr = self._correction(knl)
        beta_terms = []  # (n, coefficient) of the 2D log terms from one table
        for kind, p, coeff in split_terms(self.tree.dim, base.k, self.helmholtz_split_order):
            key = (kind, int(p))
            plan = self._term_plan(knl, key, target_boxes, starts, lists)
            tmp = torch.zeros_like(stacked)
            plan.eval(stacked[:bulk], out=tmp[:bulk], form="f64")
            if bulk < rows:
                plan.eval(stacked[bulk:], out=tmp[bulk:], form="f64")
            total = total + complex(coeff) * torch.complex(tmp[:r], tmp[r:])
            if kind == "power_log" and len(self.split_term_tables(knl, key)) == 1:
                if self.helmholtz_split_beta_mode == "table":
                    bplan = self._term_plan(knl, ("power", int(p)), target_boxes, starts, lists,
                                            log_factor=True)
                    tmp = torch.zeros_like(stacked)
                    bplan.eval(stacked[:bulk], out=tmp[:bulk], form="f64")
                    if bulk < rows:
                        bplan.eval(stacked[bulk:], out=tmp[bulk:], form="f64")
                    total = total + complex(coeff) * torch.complex(tmp[:r], tmp[r:])
                else:
                    beta_terms.append((p // 2, complex(coeff), key))
        if beta_terms:
            # one P2P for all log terms: sum_n coeff_n r^2n over the strengths
            # times log(h_s/H) (the same factor for every term: one table extent)
            n_hi = max(n for n, _, _ in beta_terms)
            pc = corr
            if n_hi > corr.nmax:  # a series long enough to hold r^2n
                from .split import HelmholtzSplitCorrection

                ck = ("beta", repr(knl), n_hi)
                if ck not in self._corrections:
                    self._corrections[ck] = HelmholtzSplitCorrection(
                        self.tree, base.k, series_terms=n_hi, deriv_kind=corr.deriv_kind,
                        deriv_axis=corr.deriv_axis)
                pc = self._corrections[ck]
            if pc.deriv_kind == _capi.SPLIT_DERIV_NONE:
                coefs = np.zeros(2 + 4 * pc.nmax)
                for n, c, _ in beta_terms:
                    coefs[2 + 4 * (n - 1) + 2] = c.real  # b_n: r^2n, no log part
                    coefs[2 + 4 * (n - 1) + 3] = c.imag
            else:  # R'(r)/r of c r^2n: 2n c r^(2n-2)
                coefs = np.zeros(4 * pc.nmax)
                for n, c, _ in beta_terms:
                    coefs[4 * (n - 1) + 2] = 2 * n * c.real
                    coefs[4 * (n - 1) + 3] = 2 * n * c.imag
            beta = self._log_alpha(self.split_term_tables(knl, beta_terms[0][2])[0], sources)
            total = total + pc(st * beta[None, :], target_boxes, starts, lists,
                               coefficients=coefs, sources=sources)
        return total[0] if one else total

    def _log_alpha(self, table, sources=None):
        """log(h_b / H) of each source's box (``_get_helmholtz_split_log_alpha_per_source``,
        4904-4974): per tree point, or per source of a smooth set; a complex
        CUDA vector."""
        key = (float(table.source_box_extent), None if sources is None else sources.m)
        if key not in self._beta:
            levels = np.asarray(self.tree.box_levels, np.int64)
            per_box = np.log(float(self.tree.root_extent) * 0.5 ** levels
                             / float(table.source_box_extent))
            if sources is not None:
                beta = sources.per_source(per_box)
            else:
                starts = np.asarray(self.tree.box_target_starts, np.int64)
                counts = np.asarray(self.tree.box_target_counts_nonchild, np.int64)
                beta = np.zeros(int(np.asarray(self.tree.points).shape[1]))
                for b in np.flatnonzero(counts > 0):
                    beta[starts[b]:starts[b] + counts[b]] = per_box[b]
            self._beta[key] = torch.as_tensor(beta, dtype=torch.complex128, device="cuda")
        return self._beta[key]

    def eval_direct_split(self, target_boxes, neighbor_source_boxes_starts,
                          neighbor_source_boxes_lists, mode_coefs, form="f64", n_slices=None,
                          fused=True, tol=1e-10):
        """List 1 as ``drive_volume_fmm`` sums it under the split
        (``volume_fmm.py:983-1026``): the Laplace-table part of
        ``mode_coefs[R, n_points]`` plus the correction with strengths
        ``mode_coefs * tree.weights``.  ``fused``: both in one plan
        contraction (``form`` "f64" or "split") where the plan takes it (2D:
        node ids in mode order); otherwise, or with ``fused=False``, the
        table part (plan ``form``) plus the point-to-point correction (and, at
        split order m > 1, the tabulated terms).  The fused plan folds split
        orders up to 17 (one table) and the smooth quadrature into its
        blocks; at order 1 it also takes per-level tables.
        ``form="split_i8"`` is accuracy-guarded on both routes as in
        ``eval_direct_plan`` (``n_slices=None``: 5 slices, then 6, then the
        FP64 form for a density whose estimate exceeds ``tol``)."""
        if not self.helmholtz_split:
            raise RuntimeError("eval_direct_split needs helmholtz_split enabled")
        self.last_split_report = []
        mc = mode_coefs if isinstance(mode_coefs, torch.Tensor) else torch.as_tensor(
            np.asarray(mode_coefs))
        mc = mc.to(device="cuda", dtype=torch.complex128)
        one = mc.dim() == 1
        mc2 = (mc.unsqueeze(0) if one else mc).contiguous()
        # the fused blocks sum a 2D log term's log(h/H) part on the quadrature
        # nodes (the "p2p" beta mode)
        beta_ok = (self.helmholtz_split_beta_mode == "p2p" or self.tree.dim != 2
                   or self.helmholtz_split_order == 1)
        # per-level Laplace tables fold in at split order 1 (the term tables of
        # higher orders are single-table)
        tables_ok = self.helmholtz_split_order == 1 or all(
            len(self.near_field_table[repr(k)]) == 1 for k in self.target_kernels)
        if (fused and beta_ok and tables_ok and not self._legacy()
                and self.helmholtz_split_order <= 17):
            # one contraction: complex blocks = table part + remainder quadrature
            # a split-form limit (slices vs the LDS budget) lets the fused
            # FP64 form (more accurate) stand in before two launches
            for fform in ((form,) if form == "f64" else (form, "f64")):
                pot = self.output_zeros(int(mc2.shape[0]))
                try:
                    for i, knl in enumerate(self.target_kernels):
                        plan = self.get_plan(knl, target_boxes, neighbor_source_boxes_starts,
                                             neighbor_source_boxes_lists, fused_split=True)
                        r = int(mc2.shape[0])
                        bulk = r - r % 64 if r > 64 else r
                        self._plan_eval(plan, mc2[:bulk], pot[i][:bulk], fform, n_slices, tol)
                        if bulk < r:
                            self._plan_eval(plan, mc2[bulk:].contiguous(), pot[i][bulk:], fform,
                                            n_slices, tol)
                        if one:
                            pot[i] = pot[i][0]
                    return self._finish(pot), TimingFuture()
                except NotImplementedError:
                    pass  # box size / node order / LDS beyond the fused plan
        pot, _ = self.eval_direct_plan(target_boxes, neighbor_source_boxes_starts,
                                       neighbor_source_boxes_lists, mc2, form=form,
                                       n_slices=n_slices, tol=tol)
        wts = torch.as_tensor(np.asarray(self.tree.weights, np.float64), device="cuda")
        corr, _ = self.eval_direct_helmholtz_split_correction(
            target_boxes, neighbor_source_boxes_starts, neighbor_source_boxes_lists,
            mc2 * wts[None, :], src_func=mc2)
        for i in range(len(pot)):
            pot[i] = pot[i] + corr[i]
            if one:
                pot[i] = pot[i][0]
        return self._finish(pot), TimingFuture()


def split_wrapper_chain(knl):
    """``(base kernel, deriv_kind, deriv_axis)`` of a split output kernel
    (``_extract_helmholtz_split_kernel_wrapper_chain``, 3884-3910): the bare
    kernel, or one axis target / source derivative around it.  Deeper
    chains raise ``NotImplementedError`` (their wrapped Laplace tables are
    nested derivatives, which this library does not build)."""
    from .kernels import AxisSourceDerivative, AxisTargetDerivative

    if isinstance(knl, (AxisTargetDerivative, AxisSourceDerivative)):
        inner = knl.inner_kernel
        if isinstance(inner, (AxisTargetDerivative, AxisSourceDerivative)):
            raise NotImplementedError(
                "helmholtz_split here supports at most one axis derivative wrapper")
        kind = (_capi.SPLIT_DERIV_TARGET if isinstance(knl, AxisTargetDerivative)
                else _capi.SPLIT_DERIV_SOURCE)
        return inner, kind, int(knl.axis)
    return knl, _capi.SPLIT_DERIV_NONE, 0


def _rewrap(knl, base):
    """``base`` under the derivative wrapper of ``knl`` (3932-3950)."""
    from .kernels import AxisSourceDerivative, AxisTargetDerivative

    if isinstance(knl, (AxisTargetDerivative, AxisSourceDerivative)):
        return type(knl)(knl.axis, base)
    return base


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


# Device-resident CSR arrays: identity key -> int32 bytes, so that a tensor
# is copied to the host (a stream sync) only the first time it is seen.  An
# entry holds its tensor, so the address cannot be reused by another tensor
# while the entry exists; the version counter changes on every in-place
# write through torch.
_DEVICE_CSR_BYTES: "OrderedDict[tuple, tuple]" = OrderedDict()
_DEVICE_CSR_MEMO_SIZE = 32


def _csr_bytes(a):
    if isinstance(a, torch.Tensor) and a.device.type != "cpu":
        ident = (a.data_ptr(), tuple(a.shape), tuple(a.stride()), str(a.dtype), str(a.device),
                 a._version)
        hit = _DEVICE_CSR_BYTES.get(ident)
        if hit is not None:
            _DEVICE_CSR_BYTES.move_to_end(ident)
            return hit[1]
        b = np.ascontiguousarray(_host(a), dtype=np.int32).tobytes()
        _DEVICE_CSR_BYTES[ident] = (a, b)
        while len(_DEVICE_CSR_BYTES) > _DEVICE_CSR_MEMO_SIZE:
            _DEVICE_CSR_BYTES.popitem(last=False)
        return b
    # host arrays: their int32 bytes (no device traffic)
    return np.ascontiguousarray(_host(a), dtype=np.int32).tobytes()


def _csr_key(out_kernel, target_boxes, starts, lists):
    """Cache key of a plan: the kernel and the int32 bytes of the CSR (for
    device tensors looked up by identity and version after the first call)."""
    return (repr(out_kernel),) + tuple(_csr_bytes(a) for a in (target_boxes, starts, lists))
