"""Device driver of the Duffy table build (``vl1_duffy_build``).

Host side of ``_batched_duffy_values_for_local_indices``
(volumential/nearfield_potential_table.py:2328-2463): gathers per-entry
geometry, uploads the node set once, launches one kernel over all entries.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi, quadrature


def _dev(a: np.ndarray, dtype=None) -> torch.Tensor:
    t = torch.from_numpy(np.array(a, copy=True, order="C"))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda", non_blocking=False)


def target_point_chunks(table, entry_ids, max_chunk=64):
    """Sort entries by target point (case, target node) and cut each run
    into near-equal chunks of at most ``max_chunk``: ``(order, starts)``
    with ``entry_ids[order]`` the sorted entries and ``starts`` the chunk
    offsets (int64, ``n_chunks + 1``)."""
    entry_ids = np.asarray(entry_ids, dtype=np.int64)
    key = (entry_ids // table.n_pairs) * table.n_q_points + entry_ids % table.n_q_points
    order = np.argsort(key, kind="stable")
    ks = key[order]
    run = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1], True])
    starts = [0]
    for a, b in zip(run[:-1], run[1:]):
        n = int(b - a)
        k = -(-n // max_chunk)
        cuts = a + (np.arange(1, k + 1) * n) // k
        starts.extend(int(c) for c in cuts)
    return order, np.asarray(starts, dtype=np.int64)


def duffy_values(table, entry_ids, regular_order, radial_order, stream=None,
                 segmented=True) -> np.ndarray:
    """Duffy values of ``entry_ids`` (one launch).  ``segmented``: share each
    node's geometry and kernel value among the entries of one target point
    (the fast form); ``False`` runs one lane per entry."""
    if not torch.cuda.is_available():
        raise RuntimeError("the Duffy table build runs on the GPU; no GPU is visible")
    L = _capi.lib()
    nodes = quadrature.duffy_node_set(table.dim, int(regular_order), int(radial_order))
    entry_ids = np.asarray(entry_ids, dtype=np.int64)
    order = None
    if segmented and entry_ids.size:
        order, starts = target_point_chunks(table, entry_ids)
        entry_ids = entry_ids[order]
        d_starts = _dev(starts)
    tp, dv, axes = table.duffy_inputs(entry_ids)
    xi = table.interp_nodes
    bw = quadrature.barycentric_weights(xi)
    n_entries = int(tp.shape[1])
    cplx = bool(table.integral_knl.is_complex_valued)
    d_tp, d_dv, d_ax = _dev(tp), _dev(dv), _dev(axes)
    d_u, d_s, d_j = _dev(nodes.node_u), _dev(nodes.node_sign), _dev(nodes.node_jac)
    d_xi, d_bw = _dev(xi), _dev(bw)
    out = torch.empty(n_entries * (2 if cplx else 1), dtype=torch.float64, device="cuda")
    import math

    nodes_per_sign = nodes.n_nodes // (2**table.dim)
    assert nodes_per_sign * 2**table.dim == nodes.n_nodes
    assert nodes_per_sign % math.factorial(table.dim) == 0
    args = _capi.DuffyArgs(
        dim=table.dim,
        quad_order=table.quad_order,
        kernel_kind=int(table.integral_knl.kind),
        helmholtz_k=float(getattr(table.integral_knl.get_base_kernel(), "k", 0.0)),
        n_entries=n_entries,
        n_nodes=nodes.n_nodes,
        nodes_per_sign=nodes_per_sign,
        source_box_extent=table.source_box_extent,
        target_points=d_tp.data_ptr(),
        decomposition=d_dv.data_ptr(),
        mode_axes=d_ax.data_ptr(),
        node_u=d_u.data_ptr(),
        node_sign=d_s.data_ptr(),
        node_jac=d_j.data_ptr(),
        interp_nodes=d_xi.data_ptr(),
        bary_w=d_bw.data_ptr(),
        result=out.data_ptr(),
        deriv_axis=int(getattr(table.integral_knl, "axis", -1)),
        chunk_start=d_starts.data_ptr() if order is not None else None,
        n_chunks=int(starts.size - 1) if order is not None else 0,
        power=int(getattr(table.integral_knl.get_base_kernel(), "power", 0)),
    )
    st = stream if stream is not None else torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(st)
    _capi.check(L.vl1_duffy_build(args, _capi.stream_handle(st)), "vl1_duffy_build")
    ev1.record(st)
    torch.cuda.synchronize()
    table.last_build_kernel_ms = ev0.elapsed_time(ev1)
    table.last_build_node_evals = int(n_entries) * int(nodes.n_nodes)
    table.last_build_chunks = int(starts.size - 1) if order is not None else int(n_entries)
    table.last_build_n_nodes = int(nodes.n_nodes)
    res = out.cpu().numpy()
    if cplx:
        res = res[0::2] + 1j * res[1::2]
    if order is not None:
        unsorted = np.empty_like(res)
        unsorted[order] = res
        res = unsorted
    return res
