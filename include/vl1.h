/*
 * vl1.h -- C ABI of the MI355X list-1 (near-field) evaluation library.
 *
 * Plain C types only: device buffers are raw pointers (allocated by the
 * caller, e.g. torch tensors), streams are opaque `void*` hipStream_t
 * handles (NULL = default stream).  No entry point throws; every one
 * returns a vl1_status and records a message for vl1_last_error().
 *
 * Reference interfaces each entry point replaces (volumential):
 *   vl1_payload_create   expansion_wrangler_fpnd.py:1805-2066,2939-3158
 *                        (_prepare_table_data_and_entry_map,
 *                         _get_cached_nearfield_payload: device upload of
 *                         table_data_combined and the arithmetic-orbit
 *                         case descriptors)
 *   vl1_near_field_f64   list1.py:1260-1487 (NearFieldFromCSR.__call__ and
 *                        the loopy `near_field` kernel, list1.py:1057-1200)
 *   vl1_plan_create /
 *   vl1_plan_eval /
 *   vl1_plan_tune        expansion_wrangler_fpnd.py:3222-3489 (eval_direct
 *                        called once per RHS by volume_fmm.py:983-989);
 *                        batched over RHS, FP64 or split-precision form
 *   vl1_plan_create_helmholtz_split
 *                        expansion_wrangler_fpnd.py:3457-3489 + 3491-3836
 *                        (table part and split correction in one plan)
 *   vl1_duffy_build      nearfield_potential_table.py:2034-2213,2328-2463
 *                        (duffy_invariant_fused_table kernel and its driver)
 *   vl1_split_correction expansion_wrangler_fpnd.py:3491-3836, 212-312,
 *                        5116-5275 (Helmholtz split correction, any split
 *                        order, base or smooth quadrature)
 *   vl1_orbit_rows_host  orbit_arithmetic.py:566-634 (host twin of the
 *                        device compact-row function, for CPU-side checks)
 */
#ifndef VL1_H
#define VL1_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VL1_ABI_VERSION 11

typedef enum vl1_status {
  VL1_OK = 0,
  VL1_E_ARG = -1,         /* invalid argument (ValueError in the reference) */
  VL1_E_LAYOUT = -2,      /* per-box node layout violation (ValueError) */
  VL1_E_LEVEL = -3,       /* single-table level mismatch (RuntimeError) */
  VL1_E_COVERAGE = -4,    /* table levels do not cover the tree (RuntimeError) */
  VL1_E_TABLE = -5,       /* missing / non-finite table data (RuntimeError) */
  VL1_E_NONFINITE = -6,   /* non-finite output (RuntimeError) */
  VL1_E_HIP = -7,         /* HIP runtime failure */
  VL1_E_UNSUPPORTED = -8  /* configuration this library does not implement */
} vl1_status;

typedef enum vl1_kernel_kind {
  VL1_KERNEL_CONSTANT = 0,  /* G = 1 */
  VL1_KERNEL_LAPLACE = 1,   /* 2D: -log|d|/(2 pi); 3D: 1/(4 pi |d|) */
  VL1_KERNEL_HELMHOLTZ = 2, /* 2D: (i/4) H0^(1)(k|d|); 3D: e^{ik|d|}/(4 pi |d|) */
  /* one axis derivative of Laplace (sumpy AxisTargetDerivative /
   * AxisSourceDerivative(axis, LaplaceKernel(dim)), table build only):
   * target dG/dd_a = 2D -d_a/(2 pi |d|^2), 3D -d_a/(4 pi |d|^3); source = minus */
  VL1_KERNEL_LAPLACE_TARGET_DERIV = 3,
  VL1_KERNEL_LAPLACE_SOURCE_DERIV = 4,
  /* one axis derivative of Helmholtz (complex, table build only): target
   * dG/dd_a = G'(|d|) d_a/|d|, G' = 2D -(ik/4) H1^(1)(k|d|), 3D
   * e^{ik|d|}(ik|d| - 1)/(4 pi |d|^2); source = minus */
  VL1_KERNEL_HELMHOLTZ_TARGET_DERIV = 5,
  VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV = 6,
  /* radial terms of the Helmholtz split beyond order 1 (real, table build
   * only; expansion_wrangler_fpnd.py:144-210): |d|^p, and |d|^p log|d|
   * (p > 0), with p = vl1_duffy_args.power */
  VL1_KERNEL_POWER = 7,
  VL1_KERNEL_POWER_LOG = 8,
  /* axis derivatives of the radial terms (deriv_axis, power p): target
   * d/dd_a |d|^p = p |d|^(p-2) d_a and d/dd_a |d|^p log|d| = d_a |d|^(p-2)
   * (p log|d| + 1); source = minus (the wrapped split terms,
   * expansion_wrangler_fpnd.py:5018-5051) */
  VL1_KERNEL_POWER_TARGET_DERIV = 9,
  VL1_KERNEL_POWER_SOURCE_DERIV = 10,
  VL1_KERNEL_POWER_LOG_TARGET_DERIV = 11,
  VL1_KERNEL_POWER_LOG_SOURCE_DERIV = 12
} vl1_kernel_kind;

/* list1.py:361-437 (get_kernel_scaling_policy) */
typedef enum vl1_scaling_mode {
  VL1_SCALING_CANONICAL = 0,  /* table level 0, scaling (h/H)^power [+ 2D log term] */
  VL1_SCALING_FIXED = 1,      /* table level 0, scaling 1 */
  VL1_SCALING_PER_LEVEL = 2,  /* table level from source level, scaling 1 */
  /* table level 0, scaling level_scaling[source level] and displacement
   * level_disp[source level] * mode_nmlz[0, s]: the caller's
   * kernel_scaling_code / kernel_displacement_code (list1.py:389-404),
   * evaluated per source-box level on the host */
  VL1_SCALING_CUSTOM = 3
} vl1_scaling_mode;
#define VL1_MAX_LEVELS 32

/* Last error message of the calling thread ("" if none). */
const char* vl1_last_error(void);
int vl1_abi_version(void);

/* ------------------------------------------------------------------ */
/* Payload: device-resident table data + orbit descriptors.            */
/* ------------------------------------------------------------------ */
typedef struct vl1_payload vl1_payload;

typedef struct vl1_payload_desc {
  int dim;            /* 2 or 3 */
  int quad_order;     /* q; n_q_points = q^dim */
  int n_cases;        /* 33 (2D) / 139 (3D) */
  int n_case_orbits;  /* canonical cases */
  int n_tables;       /* table levels */
  int is_complex;     /* 0: float64 values, 1: complex128 (interleaved re,im) */
  int64_t n_table_entries;  /* compact rows per level */
  /* host pointers, copied into the payload */
  const double* table_data;            /* [n_tables][n_table_entries] (x2 if complex) */
  const double* mode_nmlz;             /* [n_tables][n_q_points] (x2 if complex) */
  const int64_t* case_indices;         /* [13^dim] */
  const uint16_t* case_orbit_ranks;    /* [n_cases] */
  const uint8_t* case_axis_perm;       /* [n_cases][dim] */
  const int8_t* case_axis_sign;        /* [n_cases][dim] */
  const uint8_t* case_axis_group;      /* [n_cases][dim] */
  const int32_t* case_value_offsets;   /* [n_case_orbits + 1] */
  /* [dim] or NULL (all zero): arithmetic_axis_sign_power of the
   * signed-arithmetic-orbit kind (orbit_arithmetic.py:176-369,
   * list1.py:799-832).  An entry whose canonicalisation flips an axis with
   * sign power 1 reads its compact value negated. */
  const uint8_t* axis_sign_power;
} vl1_payload_desc;

int vl1_payload_create(vl1_payload** out, const vl1_payload_desc* desc, void* stream);

/* The "dense" reconstruction kind, the default of NearFieldFromCSR
 * (list1.py:120; entry reconstruction list1.py:1018-1045, has_entry gating
 * list1.py:1111-1116; arrays assembled by expansion_wrangler_fpnd.py:
 * 1887-2066 and 2980-2995).  The value of (case c, source mode s, target
 * node t) at table level lev is
 *   T[lev][ids[f]] * entry_scales[f] * mode_case_scale[s][c],
 *   f = (mode_case_map[s][c] * n + mode_qpoint_map[s][s]) * n + mode_qpoint_map[s][t],
 * and an entry with ids[f] < 0 contributes nothing, its displacement
 * included.  The payload expands every (level, case) block once on the
 * device ([n_tables][n_cases][n][n], at most 8 GiB); the arrays are
 * copied, the host pointers may be released on return. */
typedef struct vl1_dense_payload_desc {
  int dim;                /* 2 or 3 */
  int quad_order;         /* q; n = q^dim */
  int n_cases;
  int n_tables;
  int is_complex;         /* values, scales and normalisers complex128 (re, im) */
  int64_t n_table_entries;            /* layout entries per level */
  const double* table_data;           /* [n_tables][n_table_entries] (x2 if complex) */
  const double* mode_nmlz;            /* [n_tables][n] (x2 if complex) */
  const int64_t* case_indices;        /* [13^dim] */
  const int32_t* table_entry_ids;     /* [n_cases][n][n]; < 0: no entry */
  const double* table_entry_scales;   /* [n_cases][n][n] (x2 if complex) */
  const int32_t* mode_qpoint_map;     /* [n][n] */
  const int32_t* mode_case_map;       /* [n][n_cases] */
  const double* mode_case_scale;      /* [n][n_cases] (x2 if complex) */
} vl1_dense_payload_desc;

int vl1_payload_create_dense(vl1_payload** out, const vl1_dense_payload_desc* desc, void* stream);
/* 1 if a dense payload has entries without data (ids < 0) reachable from a
 * (case, s, t), else 0 (always 0 for arithmetic-orbit payloads). */
int vl1_payload_has_missing(const vl1_payload* p);
int vl1_payload_destroy(vl1_payload* p);
/* Bytes of device memory the payload holds. */
int64_t vl1_payload_device_bytes(const vl1_payload* p);

/* ------------------------------------------------------------------ */
/* NearFieldFromCSR: general CSR evaluation, reference semantics.      */
/* ------------------------------------------------------------------ */
typedef struct vl1_near_field_args {
  int scaling_mode;        /* vl1_scaling_mode */
  int scaling_power;       /* canonical: exponent of (h/H) (2 for Laplace, d for constant) */
  int log_displacement;    /* canonical 2D Laplace: add -(1/2pi) s log(h/H) nmlz */
  double root_extent;
  double table_root_extent;
  int table_starting_level;
  double encoding_shift;   /* 6 */
  int encoding_base;       /* 13 */
  double case_encoding_bias; /* 1e-10 */
  int n_tgt_boxes;
  int aligned_nboxes;
  int n_rhs;               /* number of densities (>= 1) */
  int64_t ld_coefs;        /* stride between densities, in elements */
  int64_t ld_result;       /* stride between outputs, in elements */
  /* device pointers */
  double* result;                          /* [n_rhs][ld_result] (x2 if complex) */
  const double* mode_coefs;                /* [n_rhs][ld_coefs] (x2 if complex) */
  const double* box_centers;               /* [dim][aligned_nboxes] */
  const uint8_t* box_levels;               /* [aligned_nboxes] */
  const int32_t* box_source_counts_nonchild;
  const int32_t* box_source_starts;
  const int32_t* box_target_counts_nonchild;
  const int32_t* box_target_starts;
  const int32_t* neighbor_source_boxes_starts; /* [n_tgt_boxes + 1] */
  const int32_t* neighbor_source_boxes_lists;
  const int32_t* target_boxes;             /* [n_tgt_boxes] */
  const int32_t* source_mode_ids;          /* per source particle */
  const int32_t* target_point_ids;         /* per target particle */
  int32_t* error_flags;                    /* optional [1], OR-ed: 1 = invalid case id */
  /* VL1_SCALING_CUSTOM: device arrays [VL1_MAX_LEVELS] of finite values,
   * indexed by source-box level (levels past the end take the last entry);
   * log_displacement = 0 declares level_disp all zero (no displacement pass) */
  const double* level_scaling;
  const double* level_disp;
  /* Optional (NULL = potential kind 1): potential kind 2, the inverse-type
   * (hypersingular) potentials (list1.py:101-109, 586-595, 1097-1143): each
   * target adds mode_coefs[target] x exterior_nmlz[lev][target mode] x the
   * target box's scaling (its level's table and scaling, by the rules of
   * the source boxes).  Device [n_tables][n]; real tables without a
   * displacement (VL1_E_UNSUPPORTED otherwise). */
  const double* exterior_nmlz;
} vl1_near_field_args;

/* Overwrites result[target] for every target of the listed boxes. */
int vl1_near_field_f64(const vl1_payload* p, const vl1_near_field_args* a, void* stream);

/* ------------------------------------------------------------------ */
/* Plan: batched evaluation over many densities (RHS).                 */
/* ------------------------------------------------------------------ */
typedef struct vl1_plan vl1_plan;

typedef enum vl1_eval_form {
  VL1_FORM_F64 = 0,         /* dense per-case blocks, FP64 MFMA */
  VL1_FORM_SPLIT_BF16 = 1,  /* Ozaki split: BF16 MFMA, FP32 accumulate, FP64 sums */
  VL1_FORM_SPLIT_I8 = 2     /* Ozaki split: int8 MFMA, exact int32 sums over a tile's boxes,
                               one FP64 recombination per output; error estimate per RHS */
} vl1_eval_form;

typedef struct vl1_plan_desc {
  /* host pointers describing the tree and its list-1 CSR */
  int dim;
  int n_boxes;                 /* aligned_nboxes */
  int n_tgt_boxes;
  int64_t n_points;            /* particles (tree order) */
  const double* box_centers;   /* [dim][n_boxes] */
  const uint8_t* box_levels;   /* [n_boxes] */
  const int32_t* box_starts;   /* [n_boxes] first particle of each box */
  const int32_t* box_counts;   /* [n_boxes] particles of each box */
  const int32_t* target_boxes; /* [n_tgt_boxes] */
  const int32_t* nbr_starts;   /* [n_tgt_boxes + 1] */
  const int32_t* nbr_lists;
  const int32_t* local_ids;    /* [n_points] box-local node id of each particle */
  double root_extent;
  double table_root_extent;
  int table_starting_level;
  int scaling_mode;
  int scaling_power;
  int log_displacement;
  double case_encoding_bias;
  int boxes_per_tile;          /* 0 = choose */
  /* VL1_SCALING_CUSTOM: host arrays [VL1_MAX_LEVELS] of finite values,
   * indexed by source-box level */
  const double* level_scaling;
  const double* level_disp;
} vl1_plan_desc;

typedef struct vl1_split_opts {
  int n_slices;      /* slices per operand (1..6; default 5 BF16, 6 int8) */
  int n_products;    /* 0 = all slice products with i + j < n_slices */
  /* VL1_FORM_SPLIT_I8 only, optional (NULL = none): device [n_rhs][2] that
   * receives per RHS the estimated relative error of the result over the
   * plan's targets, [0] max |error| / max |y| (Linf) and [1]
   * ||error||_2 / ||y||_2 (L2).  The estimate is 4/254 of the last slice
   * diagonal kept, the size of the first diagonal dropped, per output. */
  double* error_est;
} vl1_split_opts;

int vl1_plan_create(vl1_plan** out, const vl1_payload* p, const vl1_plan_desc* d, void* stream);

/* Helmholtz split folded into the plan (any split order, base or smooth
 * quadrature of the remainder): complex blocks per
 * (case, source level) holding the Laplace table part (canonical scaling and
 * the 2D log term folded in) plus the Gauss quadrature of the remainder
 * G_k - G_0 on the source nodes -- the table part of eval_direct
 * (expansion_wrangler_fpnd.py:3457-3489) and the correction of
 * eval_direct_helmholtz_split_correction (3491-3836) in one contraction.
 * `laplace` is a real scalar Laplace payload, d->scaling_mode canonical; the
 * plan then evaluates complex densities like a complex-table plan.  In 2D
 * the box-local node ids must be in mode order (VL1_E_UNSUPPORTED).
 * An axis target / source derivative wrapper (3884-3957) takes the wrapped
 * Laplace table's signed payload (canonical power 1, no log term) and the
 * remainder's axis derivative +-(x_a - y_a) R'(r) / r; at m > 1 the term
 * tables are the wrapped terms' (same signed layout), scaled by
 * (h/H)^(d + p - 1), and a 2D log term's part is log(h/H) x the quadrature
 * of +-(x_a - y_a) p r^(p-2) (5018-5051).  Per-level payloads
 * (d->scaling_mode per-level) are taken at split order 1. */
/* Derivative wrapper of a split output kernel (expansion_wrangler_fpnd.py:
 * 3884-3957): none, AxisTargetDerivative (d/dx_a), AxisSourceDerivative
 * (d/dy_a = -d/dx_a). */
typedef enum vl1_split_deriv {
  VL1_SPLIT_DERIV_NONE = 0,
  VL1_SPLIT_DERIV_TARGET = 1,
  VL1_SPLIT_DERIV_SOURCE = 2
} vl1_split_deriv;

typedef struct vl1_helmholtz_split_blocks {
  double k_re, k_im;          /* wave number (Yukawa: k = i lambda) */
  int series_nmax;            /* series terms, 1..96 (4337-4389) */
  const double* unit_nodes;   /* [q] Gauss-Legendre nodes on [0, 1] (host) */
  const double* unit_weights; /* [q] their weights (host) */
  const int32_t* case_vecs;   /* [n_cases][dim] list-1 gallery vectors (host) */
  int deriv_kind;             /* vl1_split_deriv: derivative wrapper of the output kernel */
  int deriv_axis;             /* its axis (deriv_kind != VL1_SPLIT_DERIV_NONE) */
  /* Split order m (0 or 1: order 1).  m > 1: the
   * remainder without the tabulated terms, plus each term's table scaled by
   * coefficient x (h/H)^(d + p) and, for a 2D log term, coefficient x
   * log(h/H) x the quadrature of r^p (expansion_wrangler_fpnd.py:3560-3836,
   * 4552-5118).  Term tables share the Laplace payload's compact layout. */
  int split_order;
  int n_terms;
  int64_t n_term_entries;       /* compact entries per term table (= the payload's) */
  const double* term_tables;    /* host [n_terms][n_term_entries] */
  const int32_t* term_powers;   /* host [n_terms] p */
  const int32_t* term_logs;     /* host [n_terms] 1: r^p log r, 0: r^p */
  const double* term_coefs;     /* host [n_terms] complex (re, im) */
  /* Smooth quadrature of the remainder (5116-5275): 0 = the base nodes;
   * m_s > 0 = per source box the m_s^d Gauss nodes, with the source mode's
   * Lagrange basis evaluated there (smooth_interp[j][i] = l_i(node_j)) */
  int smooth_order;
  const double* smooth_nodes;   /* host [smooth_order] on [0, 1] */
  const double* smooth_weights; /* host [smooth_order], summing to 1 */
  const double* smooth_interp;  /* host [smooth_order][q] */
} vl1_helmholtz_split_blocks;

int vl1_plan_create_helmholtz_split(vl1_plan** out, const vl1_payload* laplace,
                                    const vl1_plan_desc* d, const vl1_helmholtz_split_blocks* sp,
                                    void* stream);
int vl1_plan_destroy(vl1_plan* plan);
/* f: [n_rhs][ldf] densities (tree order), y: [n_rhs][ldy] outputs;
 * every target of the plan's target boxes is overwritten. */
int vl1_plan_eval(vl1_plan* plan, int form, const double* f, int64_t ldf, int n_rhs,
                  double* y, int64_t ldy, const vl1_split_opts* opts, void* stream);
/* The two halves of vl1_plan_eval: stage the densities (gather into mode
 * order, split into slices for the split form), then contract them into y.
 * vl1_plan_contract requires a preceding vl1_plan_stage with the same form,
 * RHS count and slice count on the same stream. */
int vl1_plan_stage(vl1_plan* plan, int form, const double* f, int64_t ldf, int n_rhs,
                   const vl1_split_opts* opts, void* stream);
int vl1_plan_contract(vl1_plan* plan, int form, int n_rhs, double* y, int64_t ldy,
                      const vl1_split_opts* opts, void* stream);
/* Autotuner (the reference's counterpart is the list1_target_batch_size
 * knob, list1.py / doc/source/nearfield_symmetry.rst:99-106, chosen by hand):
 * stages f, then times every valid tile-kernel variant -- bit 0: one 16-row
 * tile per wave (split form, blocks of at most 128 rows); bit 1: tiles in
 * plain Morton order instead of grouped by block signature (several boxes
 * per tile only); bit 2: no XCD phase alignment -- over `reps` contractions
 * each, and keeps the fastest for this (form, slices, RHS per tile); later
 * vl1_plan_contract / vl1_plan_eval calls use it.  y receives a valid
 * result.  Synchronous.  chosen (optional): the variant; times_ms
 * (optional, [8]): ms per contraction by variant, -1 if not valid. */
int vl1_plan_tune(vl1_plan* plan, int form, const double* f, int64_t ldf, int n_rhs, double* y,
                  int64_t ldy, const vl1_split_opts* opts, int reps, int32_t* chosen,
                  double* times_ms, void* stream);
/* Pins the tile-kernel variant (the bits of vl1_plan_tune) for this (form,
 * slices, RHS per tile) without timing, e.g. to reproduce a tuned setting
 * (list1_target_batch_size, doc/source/nearfield_symmetry.rst:99-106, is the
 * reference's hand-set counterpart); variant -1 returns to the default.
 * VL1_E_UNSUPPORTED for one row tile per wave on blocks of more than 128
 * rows.  n_slices is ignored for VL1_FORM_F64. */
int vl1_plan_set_variant(vl1_plan* plan, int form, int n_slices, int n_rhs, int variant);
/* Plan statistics: [0] list-1 pairs, [1] tiles, [2] (tile, case) steps,
 * [3] distinct (case, level) blocks, [4] device bytes. */
int vl1_plan_stats(const vl1_plan* plan, int64_t* out5);
/* Allocation-free vl1_plan_eval.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory during evaluation (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as vl1_plan_eval, but
 * the per-evaluation staging (density slices, column scales, log-term
 * dots) lives in a fixed 64 MiB device array that the library holds on
 * every GPU it runs on until the process ends; the call makes no hipMalloc
 * / hipFree.  Requirements: the plan was evaluated once with vl1_plan_eval
 * (or tuned) for this form, slice count and RHS count per tile, which
 * builds its slice set and tile schedule (VL1_E_UNSUPPORTED otherwise); the
 * staging fits the array (VL1_E_UNSUPPORTED beyond 64 MiB, e.g. 3D q = 6 at
 * 5 slices takes about 2.2 KiB per (box, RHS)).  Every call on one GPU
 * uses that GPU's array: calls that overlap in time on one GPU (other
 * streams, other host threads) race and must be serialised by the caller. */
int vl1_plan_eval_noalloc(vl1_plan* plan, int form, const double* f, int64_t ldf, int n_rhs,
                          double* y, int64_t ldy, const vl1_split_opts* opts, void* stream);

/* ------------------------------------------------------------------ */
/* Duffy table build (one launch over all canonical entries).          */
/* ------------------------------------------------------------------ */
typedef struct vl1_duffy_args {
  int dim;
  int quad_order;
  int kernel_kind;          /* vl1_kernel_kind */
  double helmholtz_k;
  int64_t n_entries;
  int64_t n_nodes;
  int64_t nodes_per_sign;   /* nodes per sign vector (perm x tail x radial) */
  double source_box_extent;
  /* device pointers */
  const double* target_points;   /* [dim][n_entries] */
  const double* decomposition;   /* [dim][n_entries] clamp(target, 0, H) */
  const int32_t* mode_axes;      /* [dim][n_entries] */
  const double* node_u;          /* [dim][n_nodes] */
  const double* node_sign;       /* [dim][n_nodes] */
  const double* node_jac;        /* [n_nodes] */
  const double* interp_nodes;    /* [quad_order] */
  const double* bary_w;          /* [quad_order] barycentric weights (validated; the kernel
                                    evaluates the basis in product form) */
  double* result;                /* [n_entries] (x2 interleaved if Helmholtz) */
  int deriv_axis;                /* derivative axis of the *_DERIV kinds */
  /* Optional (NULL = one lane per entry): entries sorted so that each chunk
   * [chunk_start[c], chunk_start[c + 1]) holds 1..64 entries of one
   * target point (same target_points and decomposition); every node's
   * geometry and kernel value is then computed once per chunk.  The
   * contract is checked on the device: a chunk outside it (more than 64 or
   * no entries, or reaching past n_entries) writes NaN to the result
   * entries it names and adds nothing else. */
  const int64_t* chunk_start;    /* device [n_chunks + 1] */
  int64_t n_chunks;
  int power;                     /* exponent p of VL1_KERNEL_POWER / _POWER_LOG (0..32) */
} vl1_duffy_args;

int vl1_duffy_build(const vl1_duffy_args* a, void* stream);

/* ------------------------------------------------------------------ */
/* Helmholtz split correction of list 1.                               */
/* Replaces eval_direct_helmholtz_split_correction,                    */
/* expansion_wrangler_fpnd.py:3491-3836, with the series remainder     */
/* kernel of 212-312 and the self limit of 5293-5317.                  */
/* ------------------------------------------------------------------ */
typedef struct vl1_split_args {
  int dim;                 /* 2 or 3 */
  int split_order;         /* m >= 1: the remainder without the terms an order-m split
                              tabulates (2D r^2n log r, 3D r^(2j-1), n, j < m;
                              expansion_wrangler_fpnd.py:284-304), also under a
                              derivative wrapper */
  int series_nmax;         /* series terms, 1..96 (expansion_wrangler_fpnd.py:4337-4389) */
  int n_tgt_boxes;
  int n_rhs;
  int max_box_points;      /* largest box_counts entry (sizes the lane layout) */
  double k_re, k_im;       /* wave number (Yukawa: k = i lambda) */
  int64_t n_points;
  /* device pointers */
  const int32_t* target_boxes;   /* [n_tgt_boxes] */
  const int32_t* nbr_starts;     /* [n_tgt_boxes + 1] list-1 CSR */
  const int32_t* nbr_lists;
  const int32_t* box_starts;     /* first point of each box (sources = targets) */
  const int32_t* box_counts;     /* points of each box */
  const double* points;          /* [dim][n_points] */
  const double* strengths;       /* [n_rhs][ld_strengths] complex (re, im) = density x weight */
  int64_t ld_strengths;          /* doubles */
  double* result;                /* [n_rhs][ld_result] complex; targets of the listed boxes
                                    overwritten */
  int64_t ld_result;             /* doubles */
  int deriv_kind;                /* vl1_split_deriv of the output kernel */
  int deriv_axis;                /* its axis */
  /* Optional (NULL = the remainder series of k, split_order and
   * series_nmax): a radial kernel given by its series coefficients
   * (host, interleaved complex), evaluated instead of the remainder and
   * at r = 0 for self pairs: 2D [c0, then (a_n, b_n) for n = 1..N],
   * value c0 + sum_n r^2n (b_n - a_n log(r) / (2 pi)); 3D [c_n for
   * n = 1..N], value sum_n c_n r^(n-1); N = series_nmax.  The 2D split of
   * order m > 1 uses it for its log(h/H) term (the r^2n part with
   * per-source log factors folded into the strengths, 3689-3800).  With a
   * derivative wrapper they give R'(r)/r instead: 2D [(a_n, b_n) for n =
   * 1..N], R'/r = sum_n r^(2n-2) (b_n - a_n log(r) / (2 pi)); 3D [c_m for
   * m = 0..N-2], R'/r = sum_m c_m r^(m-1). */
  const double* coefficients;
  /* Optional (NULL = sources are the targets): a separate source set, the
   * smooth-quadrature path (5116-5275): per box src_box_counts[b] sources
   * from src_box_starts[b] (device), coordinates src_points [dim][n_src_points];
   * strengths then index these sources ([n_rhs][ld_strengths], complex),
   * and no pair is a self pair (r = 0 takes the series value at 0, or 0
   * under a derivative wrapper). */
  const double* src_points;
  const int32_t* src_box_starts;
  const int32_t* src_box_counts;
  int64_t n_src_points;
  int max_src_box_points;        /* largest src_box_counts entry */
  /* 1: G_k - G_0 from the closed forms per pair instead of the series (the
   * reference's helmholtz_split_order1_legacy_subtraction, 3806-3822: two
   * P2P sums subtracted), self pairs taking the r = 0 limit (3824-3834), 0
   * under a derivative wrapper; split order 1, no coefficients.  2D (i/4)
   * H0^(1)(kr) + log(r) / (2 pi), H0 to 2e-12 absolute; 3D (e^{ikr} - 1) /
   * (4 pi r); derivatives d_a (G_k - G_0)'(r) / r with H1^(1) in 2D. */
  int legacy_subtraction;
} vl1_split_args;

/* corr[t] = sum over the list-1 source points s of R(|x_t - y_s|) strength[s],
 * R = G_k - G_0 as the reference's series (self pairs: its r -> 0 limit);
 * with a derivative wrapper +-(x_a - y_a) R'(r)/r (self pairs: 0, the
 * reference's limit for a wrapped chain, 5293-5300).
 * Synchronous (returns after the launch completes). */
int vl1_split_correction(const vl1_split_args* a, void* stream);

/* ------------------------------------------------------------------ */
/* List-1 CSR of a 2:1-balanced tree of leaves, built on the GPU.      */
/* Replaces boxtree's list-1 traversal used by meshgen.py:391-444 /    */
/* tree_interactive_build.py:202-470 for the leaf-only trees here.     */
/* ------------------------------------------------------------------ */
typedef struct vl1_csr_builder vl1_csr_builder;
/* levels [n_leaves] and index [n_leaves][dim] (integer cell index of each
 * leaf at its level) are HOST arrays.  Sorts the leaves and counts every
 * leaf's neighbours (leaves touching it, itself included) on the device;
 * *n_pairs receives the total.  VL1_E_LAYOUT if the leaves are not a
 * 2:1-balanced tiling of the cube. */
int vl1_csr_builder_create(vl1_csr_builder** out, int dim, int n_leaves, const uint8_t* levels,
                           const int32_t* index, int64_t* n_pairs, void* stream);
/* Writes starts [n_leaves + 1] and lists [n_pairs] (device, caller-owned);
 * each list sorted by leaf id. */
int vl1_csr_builder_fill(vl1_csr_builder* b, int32_t* starts, int32_t* lists, void* stream);
int vl1_csr_builder_destroy(vl1_csr_builder* b);

/* Host evaluation of the compact-row function the device kernels use
 * (orbit_arithmetic.py:566-634 _evaluate_compact_arithmetic_orbit_entry,
 * row relative to the orbit's first row): out[i] = row(s[i], t[i]) for one
 * case's descriptors perm/sign/group [dim]. */
int vl1_orbit_rows_host(int dim, int q, const uint8_t* perm, const int8_t* sign,
                        const uint8_t* group, const int32_t* s, const int32_t* t,
                        int64_t count, int64_t* out);
/* Host evaluation of the entry sign the device kernels apply for signed
 * (axis-derivative) kernels (orbit_arithmetic.py:372-432 entry_sign):
 * out[i] = +1 / -1 for (s[i], t[i]) of one case, axis_sign_power [dim]. */
int vl1_orbit_signs_host(int dim, int q, const uint8_t* perm, const int8_t* sign,
                         const uint8_t* axis_sign_power, const int32_t* s, const int32_t* t,
                         int64_t count, int8_t* out);

#ifdef __cplusplus
}
#endif

#endif /* VL1_H */
