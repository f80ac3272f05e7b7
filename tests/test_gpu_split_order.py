"""Helmholtz split of order m > 1 (expansion_wrangler_fpnd.py:212-312,
3560-3836, 4552-5118) on the GPU.

* Parity: list 1 at order m (the Laplace-table part + the order-m remainder
  over list 1 + each tabulated term's table evaluation times its
  coefficient, and in 2D the log(h/H) part of the log terms) matches the
  oracle assembled from the same pieces -- the numpy list-1 oracle (custom
  scaling (h/H)^(d+p) on the GPU-built term tables, whose entries
  test_gpu_duffy pins) and the P2P oracle with the order-m series and the
  r^2n kernel -- to 1e-12 (rel L2 and Linf per RHS), orders 2 and 3, 2D
  and 3D.
  Both routes: the two-launch one (table plans + point-to-point sums) and
  the fused plan, whose complex blocks carry the order-m remainder, the
  term tables and the log parts (FP64 form 1e-12, int8 split form 1e-10).
* The smooth-quadrature path (``helmholtz_split_smooth_quad_order`` > q):
  the remainder over the density interpolated to finer per-box Gauss nodes
  matches the oracle's restatement to 1e-12, orders 1 and 2, both routes
  (the fused int8 split form to 1e-10).
* Physics: against list 1 evaluated with directly built Helmholtz tables,
  order 2 is closer than order 1 (the tabulated terms are the non-smooth
  ones the base quadrature integrates least well), and so is "auto".
"""

import numpy as np
import pytest
import torch

from oracle import list1_ref, orbit_ref
from oracle import split_ref as S
from tests.gpu_helpers import adaptive, rel_errors
from volumential_amd import densities, tree as vtree
from volumential_amd.kernels import HelmholtzKernel, LaplaceKernel
from volumential_amd.table import DuffyBuildConfig, NearFieldInteractionTable
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu

RULE = DuffyBuildConfig(regular_quad_order=14, radial_quad_order=41)


def _setup(dim, q, level, k, m, R=3):
    tr = vtree.uniform_tree(dim, q, level)
    knl = HelmholtzKernel(dim, k)
    lap = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=2.0)
    lap.build_table(RULE)
    w = NearFieldWrangler(tr, {repr(knl): [lap]}, target_kernels=[knl], helmholtz_split_order=m,
                          helmholtz_split_term_build_config=RULE)
    f = np.stack([densities.helmholtz_beams(tr.points, 70 + r, k=k) for r in range(R)])
    return tr, knl, lap, w, f


def _oracle(tr, knl, lap, w, f, m, smooth=None):
    dim = tr.dim
    meta = orbit_ref.case_metadata(dim, tr.quad_order)
    kw = dict(table_root_extent=2.0, own_csr=False)

    def table_part(tab, **mode):
        out = [list1_ref.near_field_from_tree(tr, tab.compact_values()[None],
                                              tab.mode_normalizers[None], meta, part, **mode, **kw)
               for part in (f.real, f.imag)]
        return out[0] + 1j * out[1]

    tot = table_part(lap, mode="canonical", scaling_power=2, log_displacement=dim == 2)
    wts = np.asarray(tr.weights)[None, :]
    corr = w._correction(knl)
    if smooth is None:
        st, src = f * wts, None
        beta_of = lambda per_box: per_box[_box_of_point(tr)]
    else:
        pts, starts, counts, st = S.smooth_sources(tr, f, smooth)
        src = (pts, starts, counts)
        beta_of = lambda per_box: np.repeat(per_box, counts)
    tot = tot + S.correction(tr, st, knl.k, corr.nmax, split_order=m, sources=src)
    h = tr.root_extent * 0.5 ** np.arange(32)
    for kind, p, c in S.extracted_terms(dim, knl.k, m):
        tt = w.split_term_tables(knl, (kind, p))[0]
        part = table_part(tt, mode="custom", level_scaling=(h / 2.0) ** (dim + p),
                          level_disp=np.zeros(32))
        if kind == "power_log":
            per_box = np.log(tr.root_extent * 0.5 ** np.asarray(tr.box_levels, float) / 2.0)
            beta = beta_of(per_box)
            part = part + S.correction(tr, st * beta[None, :], knl.k, corr.nmax,
                                       kernel=lambda r, p=p: r ** p, sources=src)
        tot = tot + c * part
    return tot


def _box_of_point(tr):
    out = np.zeros(np.asarray(tr.points).shape[1], np.int64)
    for b in np.flatnonzero(np.asarray(tr.box_target_counts_nonchild) > 0):
        out[tr.box_target_starts[b]:tr.box_target_starts[b] + tr.box_target_counts_nonchild[b]] = b
    return out


ROUTES = [(False, "f64", 1e-12), (True, "f64", 1e-12), (True, "split_i8", 1e-10)]


@pytest.mark.parametrize("fused,form,tol", ROUTES)
@pytest.mark.parametrize("dim,q,level,k,m,ms", [(2, 4, 3, 8.0, 1, 6), (2, 4, 3, 8.0, 2, 7),
                                                (3, 3, 2, 6.0, 1, 5), (3, 3, 2, 6.0, 2, 4)])
def test_smooth_quadrature_path_matches_oracle(cuda, dim, q, level, k, m, ms, fused, form, tol):
    """helmholtz_split_smooth_quad_order > q: the remainder (and the 2D
    log(h/H) term) over the interpolated density on the finer nodes."""
    print(f"case: smooth quadrature order {ms}, split order {m}, {dim}D q={q} k={k}, "
          f"fused={fused} {form}", flush=True)
    tr = vtree.uniform_tree(dim, q, level)
    knl = HelmholtzKernel(dim, k)
    lap = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=2.0)
    lap.build_table(RULE)
    w = NearFieldWrangler(tr, {repr(knl): [lap]}, target_kernels=[knl], helmholtz_split_order=m,
                          helmholtz_split_term_build_config=RULE,
                          helmholtz_split_smooth_quad_order=ms)
    f = np.stack([densities.helmholtz_beams(tr.points, 80 + r, k=k, width=0.3) for r in range(2)])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_split(*args, torch.as_tensor(f, device="cuda"), form=form,
                              fused=fused)[0][0]
    ref = _oracle(tr, knl, lap, w, f, m, smooth=ms)
    l2, linf = rel_errors(got.cpu().numpy(), ref)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol


@pytest.mark.parametrize("fused,form,tol", ROUTES)
@pytest.mark.parametrize("dim,q,level,k,m", [(3, 3, 2, 6.0, 2), (3, 3, 2, 6.0, 3),
                                             (2, 4, 3, 12.0, 2), (2, 4, 3, 12.0, 3),
                                             (3, 5, 2, 4.0, 2)])
def test_order_m_matches_oracle(cuda, dim, q, level, k, m, fused, form, tol):
    """3D q = 5: complex blocks of 250 rows, cut into two row groups."""
    print(f"case: split order {m}, {dim}D q={q} level {level} k={k}, fused={fused} {form}",
          flush=True)
    tr, knl, lap, w, f = _setup(dim, q, level, k, m)
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_split(*args, torch.as_tensor(f, device="cuda"), form=form,
                              fused=fused)[0][0]
    ref = _oracle(tr, knl, lap, w, f, m)
    l2, linf = rel_errors(got.cpu().numpy(), ref)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol
    if fused:  # the fused plan was taken (no fall-back to two launches)
        key = [k for k in w._plans if k[-1]]
        assert key, "the fused split plan was not built"


@pytest.mark.parametrize("fused,form,tol", ROUTES)
@pytest.mark.parametrize("dim,q,max_level,k,m,ms", [(2, 4, 4, 8.0, 3, None), (2, 4, 4, 8.0, 2, 6),
                                                    (3, 3, 4, 5.0, 2, None), (3, 3, 4, 5.0, 2, 5)])
def test_adaptive_trees_match_oracle(cuda, dim, q, max_level, k, m, ms, fused, form, tol):
    """Source boxes on several levels: the term scaling (h/H)^(d+p), the 2D
    log(h/H) factor and the remainder geometry change with the source
    level (the fused plan's blocks are per (case, source level))."""
    print(f"case: adaptive {dim}D q={q} max level {max_level} k={k}, split order {m}, "
          f"smooth {ms}, fused={fused} {form}", flush=True)
    tr = adaptive(dim, q, max_level=max_level)
    assert len(np.unique(np.asarray(tr.box_levels)[tr.neighbor_source_boxes_lists])) >= 3
    knl = HelmholtzKernel(dim, k)
    lap = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=2.0)
    lap.build_table(RULE)
    w = NearFieldWrangler(tr, {repr(knl): [lap]}, target_kernels=[knl], helmholtz_split_order=m,
                          helmholtz_split_term_build_config=RULE,
                          helmholtz_split_smooth_quad_order=ms)
    f = np.stack([densities.helmholtz_beams(tr.points, 60 + r, k=k, width=0.3) for r in range(3)])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_split(*args, torch.as_tensor(f, device="cuda"), form=form,
                              fused=fused)[0][0]
    ref = _oracle(tr, knl, lap, w, f, m, smooth=ms)
    l2, linf = rel_errors(got.cpu().numpy(), ref)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol


@pytest.mark.parametrize("form,tol", [("f64", 1e-12), ("split_i8", 1e-10)])
@pytest.mark.parametrize("ms", [None, 6])
def test_fused_per_level_tables_match_oracle(cuda, ms, form, tol):
    """Split order 1 over per-level Laplace tables (no scaling, no log term):
    the fused plan's blocks read each source level's own table, on the base
    or the smooth quadrature, vs the oracle's per-level list 1 plus its
    remainder sum."""
    print(f"case: fused split over per-level tables, adaptive 2D q=4, smooth {ms}, {form}",
          flush=True)
    from tests.gpu_helpers import synthetic_table_levels

    tr = adaptive(2, 4)
    tables = synthetic_table_levels(2, 4, n_tables=tr.nlevels, seed=9)
    k = 6.0 + 0.5j
    knl = HelmholtzKernel(2, k)
    w = NearFieldWrangler(tr, {repr(knl): tables}, target_kernels=[knl],
                          helmholtz_split_smooth_quad_order=ms)
    f = np.stack([densities.helmholtz_beams(tr.points, 50 + r, k=6.0, width=0.3)
                  for r in range(3)])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_split(*args, torch.as_tensor(f, device="cuda"), form=form)[0][0]
    assert any(key[-1] for key in w._plans), "the fused plan was not built"
    meta = orbit_ref.case_metadata(2, 4)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    part = [list1_ref.near_field_from_tree(
        tr, data, nm, meta, comp, mode="per_level", table_root_extent=2.0,
        table_starting_level=0, own_csr=False) for comp in (f.real, f.imag)]
    nmax = w._correction(knl).nmax
    if ms is None:
        corr = S.correction(tr, f * np.asarray(tr.weights)[None, :], k, nmax)
    else:
        pts, starts, counts, st = S.smooth_sources(tr, f, ms)
        corr = S.correction(tr, st, k, nmax, sources=(pts, starts, counts))
    ref = part[0] + 1j * part[1] + corr
    l2, linf = rel_errors(got.cpu().numpy(), ref)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol


@pytest.mark.parametrize("tree_kind,m", [("uniform", 2), ("uniform", 3), ("adaptive", 3)])
def test_table_beta_mode_matches_oracle(cuda, tree_kind, m):
    """power_log_single_table_beta_mode="table" (3757-3779): the log(h/H)
    part of each 2D r^2n log r term is the r^2n term table on the density
    times log(h_s/H), here against the oracle's custom-scaling list 1 over
    the same r^2n table (1e-12).  At q = 4 the base quadrature integrates
    r^2n times a basis polynomial exactly for n <= 2, so the default "p2p"
    mode agrees with it to the tables' build accuracy (1e-10); the fused
    plan, which sums the p2p form, is not taken."""
    dim, q, k = 2, 4, 8.0
    print(f"case: table beta mode, 2D q={q} {tree_kind} k={k}, split order {m}", flush=True)
    tr = vtree.uniform_tree(dim, q, 3) if tree_kind == "uniform" else adaptive(dim, q)
    knl = HelmholtzKernel(dim, k)
    lap = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=2.0)
    lap.build_table(RULE)
    kw = dict(target_kernels=[knl], helmholtz_split_order=m,
              helmholtz_split_term_build_config=RULE)
    w = NearFieldWrangler(tr, {repr(knl): [lap]}, helmholtz_split_auto_config={
        "power_log_single_table_beta_mode": "table"}, **kw)
    f = np.stack([densities.helmholtz_beams(tr.points, 40 + r, k=k, width=0.3) for r in range(2)])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_split(*args, torch.as_tensor(f, device="cuda"))[0][0].cpu().numpy()
    assert not any(key[-1] for key in w._plans), "the fused plan sums the p2p form"
    # oracle: the p2p oracle minus its log(h/H) P2P parts plus the table forms
    meta = orbit_ref.case_metadata(dim, q)
    kwo = dict(table_root_extent=2.0, own_csr=False)
    tot = _oracle(tr, knl, lap, w, f, m)
    wts = np.asarray(tr.weights)[None, :]
    corr = w._correction(knl)
    h = tr.root_extent * 0.5 ** np.arange(32)
    per_box = np.log(tr.root_extent * 0.5 ** np.asarray(tr.box_levels, float) / 2.0)
    beta = per_box[_box_of_point(tr)]
    for kind, p, c in S.extracted_terms(dim, knl.k, m):
        tot = tot - c * S.correction(tr, f * wts * beta[None, :], knl.k, corr.nmax,
                                     kernel=lambda r, p=p: r ** p)
        pt = w.split_term_tables(knl, ("power", p))[0]
        part = [list1_ref.near_field_from_tree(
            tr, pt.compact_values()[None], pt.mode_normalizers[None], meta, comp, mode="custom",
            level_scaling=(h / 2.0) ** (dim + p) * np.log(h / 2.0), level_disp=np.zeros(32),
            **kwo) for comp in (f.real, f.imag)]
        tot = tot + c * (part[0] + 1j * part[1])
    l2, linf = rel_errors(got, tot)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= 1e-12 and linf <= 1e-12
    wp = NearFieldWrangler(tr, {repr(knl): [lap]}, **kw)
    p2p = wp.eval_direct_split(*args, torch.as_tensor(f, device="cuda"), fused=False)[0][0]
    l2, linf = rel_errors(got, p2p.cpu().numpy())
    print(f"vs the p2p mode: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= 1e-10 and linf <= 1e-10


@pytest.mark.parametrize("dim,q,level,k", [(3, 4, 2, 10.0), (2, 4, 3, 8.0)])
def test_order_2_closer_to_helmholtz_tables(cuda, dim, q, level, k):
    print(f"case: split orders 1-3 vs Helmholtz tables, {dim}D q={q} level {level} k={k}",
          flush=True)
    tr = vtree.uniform_tree(dim, q, level)
    knl = HelmholtzKernel(dim, k)
    ext = tr.root_extent / 2 ** level
    helm = NearFieldInteractionTable(q, dim, knl, source_box_extent=ext, source_box_level=level)
    helm.build_table(DuffyBuildConfig(regular_quad_order=20, radial_quad_order=61))
    f = torch.as_tensor(np.stack([densities.helmholtz_beams(tr.points, 90 + r, k=k, width=0.3)
                                  for r in range(2)]), device="cuda")
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    wd = NearFieldWrangler(tr, {repr(knl): [helm]}, target_kernels=[knl], helmholtz_split=False)
    direct = wd.eval_direct_batched(*args, f)[0][0].cpu().numpy()
    lap = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=2.0)
    lap.build_table(DuffyBuildConfig(regular_quad_order=20, radial_quad_order=61))
    errs = []
    for m in (1, 2, 3):
        w = NearFieldWrangler(tr, {repr(knl): [lap]}, target_kernels=[knl], helmholtz_split_order=m)
        got = w.eval_direct_split(*args, f, fused=False)[0][0].cpu().numpy()
        errs.append(rel_errors(got, direct)[0])
    print("rel L2 vs Helmholtz tables, orders 1-3:", ", ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert errs[1] < errs[0]
    # "auto": the order from |k| h and, above q, the smooth quadrature
    wa = NearFieldWrangler(tr, {repr(knl): [lap]}, target_kernels=[knl],
                           helmholtz_split_order="auto")
    got = wa.eval_direct_split(*args, f)[0][0].cpu().numpy()
    print(f"auto: order {wa.helmholtz_split_order}, smooth order "
          f"{wa.helmholtz_split_smooth_quad_order}, rel L2 {rel_errors(got, direct)[0]:.2e}",
          flush=True)
    assert rel_errors(got, direct)[0] < errs[0]


@pytest.mark.parametrize("form,tol", [("f64", 1e-12), ("split_i8", 1e-10)])
@pytest.mark.parametrize("dim,q,tree_kind,wrap,axis,k,m", [
    (2, 4, "uniform", "target", 1, 8.0, 2), (2, 4, "adaptive", "source", 0, 6.0 + 0.5j, 3),
    (3, 3, "uniform", "target", 2, 5.0, 2), (3, 3, "uniform", "source", 0, 4.0 + 0.5j, 3)])
def test_derivative_wrapper_order_m_matches_oracle(cuda, dim, q, tree_kind, wrap, axis, k, m,
                                                   form, tol):
    """Split order m > 1 under an axis derivative wrapper (5018-5051): the
    wrapped Laplace table (power 1) + the derivative of the order-m
    remainder + each wrapped term table scaled by (h/H)^(d+p-1) + in 2D the
    log(h/H) part as the derivative P2P of r^2n -- vs the oracle assembled
    from the same pieces."""
    from oracle import cpu as ocpu
    from volumential_amd.kernels import AxisSourceDerivative, AxisTargetDerivative

    print(f"case: {wrap} derivative axis {axis}, split order {m}, {dim}D q={q} {tree_kind} "
          f"k={k} {form}", flush=True)
    tr = (vtree.uniform_tree(dim, q, 3 if dim == 2 else 2) if tree_kind == "uniform"
          else adaptive(dim, q))
    cls = AxisTargetDerivative if wrap == "target" else AxisSourceDerivative
    sgn = 1 if wrap == "target" else -1
    hk = cls(axis, HelmholtzKernel(dim, k))
    dtab = NearFieldInteractionTable(q, dim, cls(axis, LaplaceKernel(dim)), source_box_extent=2.0)
    dtab.build_table(RULE)
    w = NearFieldWrangler(tr, {repr(hk): [dtab]}, target_kernels=[hk], helmholtz_split_order=m,
                          helmholtz_split_term_build_config=RULE)
    f = np.stack([densities.helmholtz_beams(tr.points, 30 + r, k=4.0, width=0.3)
                  for r in range(2)])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    fd = torch.as_tensor(f, device="cuda")
    got = w.eval_direct_split(*args, fd, form=form)[0][0]
    assert any(key[-1] for key in w._plans), "the fused plan was not built"
    two = w.eval_direct_split(*args, fd, form=form, fused=False)[0][0]
    power = [0] * dim
    power[axis] = 1
    meta = orbit_ref.case_metadata(dim, q, (axis,), power)
    part = ocpu.near_field(tr, dtab.compact_values(), dtab.mode_normalizers, meta,
                           np.concatenate([f.real, f.imag]), mode="canonical", scaling_power=1,
                           table_root_extent=2.0)
    ref = part[:2] + 1j * part[2:]
    st = f * np.asarray(tr.weights)[None, :]
    nmax = w._correction(hk).nmax
    ref = ref + S.correction(tr, st, k, nmax, deriv_axis=axis, deriv_sign=sgn, split_order=m)
    h = tr.root_extent * 0.5 ** np.arange(32)
    per_box = np.log(tr.root_extent * 0.5 ** np.asarray(tr.box_levels, float) / 2.0)
    beta = per_box[_box_of_point(tr)]
    for kind, p, c in S.extracted_terms(dim, k, m):
        tt = w.split_term_tables(hk, (kind, p))[0]
        tp = [list1_ref.near_field_from_tree(
            tr, tt.compact_values()[None], tt.mode_normalizers[None], meta, comp, mode="custom",
            level_scaling=(h / 2.0) ** (dim + p - 1), level_disp=np.zeros(32),
            table_root_extent=2.0, own_csr=False) for comp in (f.real, f.imag)]
        term = tp[0] + 1j * tp[1]
        if kind == "power_log":
            term = term + S.correction(tr, st * beta[None, :], k, nmax, deriv_axis=axis,
                                       deriv_sign=sgn,
                                       kernel_derivative=lambda r, p=p: p * r ** (p - 2.0))
        ref = ref + c * term
    for route, val in (("fused", got), ("two launches", two)):
        l2, linf = rel_errors(val.cpu().numpy(), ref)
        print(f"{route} vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
        assert l2 <= tol and linf <= tol


@pytest.mark.parametrize("dim,q,level,k", [(3, 4, 2, 10.0), (2, 4, 3, 8.0)])
def test_derivative_order_2_closer_to_helmholtz_tables(cuda, dim, q, level, k):
    """Physics of the wrapped split: against list 1 through directly built
    Helmholtz target-derivative tables, order 2 is closer than order 1 (a
    wrong term scaling power or sign would not be)."""
    from volumential_amd.kernels import AxisTargetDerivative

    print(f"case: derivative split orders 1-3 vs Helmholtz derivative tables, {dim}D q={q}",
          flush=True)
    tr = vtree.uniform_tree(dim, q, level)
    axis = dim - 1
    hk = AxisTargetDerivative(axis, HelmholtzKernel(dim, k))
    ext = tr.root_extent / 2 ** level
    helm = NearFieldInteractionTable(q, dim, hk, source_box_extent=ext, source_box_level=level)
    helm.build_table(DuffyBuildConfig(regular_quad_order=20, radial_quad_order=61))
    f = torch.as_tensor(np.stack([densities.helmholtz_beams(tr.points, 95 + r, k=k, width=0.3)
                                  for r in range(2)]), device="cuda")
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    wd = NearFieldWrangler(tr, {repr(hk): [helm]}, target_kernels=[hk], helmholtz_split=False)
    direct = wd.eval_direct_batched(*args, f)[0][0].cpu().numpy()
    lap = NearFieldInteractionTable(q, dim, AxisTargetDerivative(axis, LaplaceKernel(dim)),
                                    source_box_extent=2.0)
    lap.build_table(DuffyBuildConfig(regular_quad_order=20, radial_quad_order=61))
    errs = []
    for m in (1, 2, 3):
        w = NearFieldWrangler(tr, {repr(hk): [lap]}, target_kernels=[hk], helmholtz_split_order=m)
        got = w.eval_direct_split(*args, f, fused=False)[0][0].cpu().numpy()
        errs.append(rel_errors(got, direct)[0])
    print("rel L2 vs Helmholtz derivative tables, orders 1-3:",
          ", ".join(f"{e:.2e}" for e in errs), flush=True)
    assert errs[1] < errs[0] and errs[2] < errs[0]


@pytest.mark.parametrize("m", [1, 2])
def test_derivative_wrapper_smooth_quadrature_matches_oracle(cuda, m):
    """The smooth-quadrature path under a derivative wrapper: the derivative
    remainder (and at m = 2 the log part's derivative P2P) over the
    interpolated density on finer nodes, coincident node pairs giving 0
    (the wrapped chain's limit), vs the oracle."""
    from oracle import cpu as ocpu
    from volumential_amd.kernels import AxisTargetDerivative

    dim, q, k, ms, axis = 2, 4, 6.0, 7, 0
    print(f"case: target derivative, smooth order {ms}, split order {m}, 2D q={q}", flush=True)
    tr = vtree.uniform_tree(dim, q, 3)
    hk = AxisTargetDerivative(axis, HelmholtzKernel(dim, k))
    dtab = NearFieldInteractionTable(q, dim, AxisTargetDerivative(axis, LaplaceKernel(dim)),
                                     source_box_extent=2.0)
    dtab.build_table(RULE)
    w = NearFieldWrangler(tr, {repr(hk): [dtab]}, target_kernels=[hk], helmholtz_split_order=m,
                          helmholtz_split_term_build_config=RULE,
                          helmholtz_split_smooth_quad_order=ms)
    f = np.stack([densities.helmholtz_beams(tr.points, 60 + r, k=k, width=0.3) for r in range(2)])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_split(*args, torch.as_tensor(f, device="cuda"))[0][0].cpu().numpy()
    assert any(key[-1] for key in w._plans), "the fused plan was not built"
    power = [0] * dim
    power[axis] = 1
    meta = orbit_ref.case_metadata(dim, q, (axis,), power)
    part = ocpu.near_field(tr, dtab.compact_values(), dtab.mode_normalizers, meta,
                           np.concatenate([f.real, f.imag]), mode="canonical", scaling_power=1,
                           table_root_extent=2.0)
    ref = part[:2] + 1j * part[2:]
    pts, starts, counts, st = S.smooth_sources(tr, f, ms)
    src = (pts, starts, counts)
    nmax = w._correction(hk).nmax
    ref = ref + S.correction(tr, st, k, nmax, deriv_axis=axis, split_order=m, sources=src)
    h = tr.root_extent * 0.5 ** np.arange(32)
    beta = np.repeat(np.log(tr.root_extent * 0.5 ** np.asarray(tr.box_levels, float) / 2.0),
                     counts)
    for kind, p, c in S.extracted_terms(dim, k, m):
        tt = w.split_term_tables(hk, (kind, p))[0]
        tp = [list1_ref.near_field_from_tree(
            tr, tt.compact_values()[None], tt.mode_normalizers[None], meta, comp, mode="custom",
            level_scaling=(h / 2.0) ** (dim + p - 1), level_disp=np.zeros(32),
            table_root_extent=2.0, own_csr=False) for comp in (f.real, f.imag)]
        term = tp[0] + 1j * tp[1] + S.correction(
            tr, st * beta[None, :], k, nmax, deriv_axis=axis, sources=src,
            kernel_derivative=lambda r, p=p: p * r ** (p - 2.0))
        ref = ref + c * term
    l2, linf = rel_errors(got, ref)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= 1e-12 and linf <= 1e-12


def test_derivative_wrapper_table_beta_mode(cuda):
    """The "table" beta mode under a derivative wrapper: the log(h/H) part
    of each 2D r^2n log r term through the wrapped r^2n table (scaled by
    (h/H)^(d+2n-1) log(h/H)).  At q = 4 the base quadrature of d_a r^2n
    times a basis polynomial is exact for n <= 2, so it equals the p2p
    mode to the tables' build accuracy."""
    from volumential_amd.kernels import AxisTargetDerivative

    print("case: table beta mode under a target derivative, 2D q=4 adaptive, order 3",
          flush=True)
    tr = adaptive(2, 4)
    hk = AxisTargetDerivative(1, HelmholtzKernel(2, 6.0))
    dtab = NearFieldInteractionTable(4, 2, AxisTargetDerivative(1, LaplaceKernel(2)),
                                     source_box_extent=2.0)
    dtab.build_table(RULE)
    kw = dict(target_kernels=[hk], helmholtz_split_order=3, helmholtz_split_term_build_config=RULE)
    f = torch.as_tensor(np.stack([densities.helmholtz_beams(tr.points, 11 + r, k=6.0, width=0.3)
                                  for r in range(2)]), device="cuda")
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    wt = NearFieldWrangler(tr, {repr(hk): [dtab]}, helmholtz_split_auto_config={
        "power_log_single_table_beta_mode": "table"}, **kw)
    got = wt.eval_direct_split(*args, f)[0][0].cpu().numpy()
    wp = NearFieldWrangler(tr, {repr(hk): [dtab]}, **kw)
    p2p = wp.eval_direct_split(*args, f, fused=False)[0][0].cpu().numpy()
    l2, linf = rel_errors(got, p2p)
    print(f"table vs p2p mode: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= 1e-10 and linf <= 1e-10
