"""Oracle: Duffy quadrature of near-field table entries (numpy).

Restates:
* ``singular_integral_2d.py:807-853`` -- GL tail rule and the closed-form
  "tanh-sinh-fast" radial rule;
* ``singular_integral_2d.py:856-949`` -- ``box_quad_duffy_radial_nd``: split
  the box at the (clamped) singular point into 2^d sub-boxes x d! simplices,
  Duffy-map each to the unit cube;
* ``singular_integral_2d.py:594-761`` -- the 2D triangle form (4 triangles
  at the singular point, GL in theta on [0, pi/4], radial rule in rho);
* ``nearfield_potential_table.py:917-946`` (``find_target_point``),
  ``749-801`` (``get_mode``: tensor barycentric Lagrange modes) and
  ``1099-1177`` (``compute_table_entry_duffy_radial``);
* ``lagrange.py:45-92`` -- barycentric weights and basis evaluation.

Closed forms used to pin the quadrature (independent of the reference):
``box_potential_laplace3d`` (Newtonian potential of a uniform box),
``box_potential_laplace2d`` (logarithmic potential of a uniform rectangle)
and their gradients (``box_potential_gradient``, for the axis-derivative
kernels).
"""

from __future__ import annotations

import itertools

import numpy as np


# ---------------------------------------------------------------- rules
def gl01(order):
    x, w = np.polynomial.legendre.leggauss(int(order))
    return 0.5 * (x + 1.0), 0.5 * w


def tanh_sinh_fast(order):
    n = max(3, int(order))
    h = 1.0 / np.sqrt(n)
    k = np.arange(-n, n + 1, dtype=np.float64)
    t = h * k
    sh, ch = np.sinh(t), np.cosh(t)
    arg = 0.5 * np.pi * sh
    rho = 0.5 * (1.0 + np.tanh(arg))
    e = np.exp(-2.0 * np.abs(arg))
    sech2 = 4.0 * e / (1.0 + e) ** 2
    w = 0.25 * np.pi * h * ch * sech2
    m = (rho > 0.0) & (rho < 1.0) & np.isfinite(w)
    return rho[m], w[m]


def tail_nodes(dim_minus_one, order):
    if dim_minus_one == 0:
        return [np.array([])], [1.0]
    x, w = gl01(order)
    gx = np.meshgrid(*([x] * dim_minus_one), indexing="ij")
    gw = np.meshgrid(*([w] * dim_minus_one), indexing="ij")
    return np.stack([g.ravel() for g in gx], axis=-1), np.prod(np.array(gw), axis=0).ravel()


# -------------------------------------------------------------- lagrange
def bary_weights(nodes):
    nodes = np.asarray(nodes, np.float64)
    d = nodes[:, None] - nodes[None, :]
    np.fill_diagonal(d, 1.0)
    s = np.prod(np.sign(d), axis=1)
    lw = -np.sum(np.log(np.abs(d)), axis=1)
    return s * np.exp(lw - np.max(lw))


def lagrange_1d(nodes, index, x, weights):
    """Exact at nodes (``diff == 0``), barycentric elsewhere."""
    x = np.asarray(x, np.float64)
    diffs = x[..., None] - nodes
    flat = diffs.reshape(-1, nodes.size)
    exact = flat == 0.0
    row_exact = np.any(exact, axis=1)
    out = np.empty(flat.shape[0])
    out[row_exact] = exact[row_exact, index]
    if np.any(~row_exact):
        terms = weights / flat[~row_exact]
        out[~row_exact] = terms[:, index] / np.sum(terms, axis=1)
    return out.reshape(x.shape)


# --------------------------------------------------------------- kernels
def laplace(dim):
    if dim == 2:
        return lambda *d: -0.25 / np.pi * np.log(sum(c * c for c in d))
    return lambda *d: 0.25 / np.pi / np.sqrt(sum(c * c for c in d))


def laplace_target_derivative(dim, axis):
    """sumpy ``AxisTargetDerivative(axis, LaplaceKernel(dim))`` with its
    global scaling: ``dG/dd_axis`` (d = target - source) of ``laplace(dim)``
    -- 2D ``-d_a / (2 pi r^2)``, 3D ``-d_a / (4 pi r^3)``.  (sumpy is not
    importable here; the formula is the derivative of ``laplace``, pinned by
    finite differences of the closed-form box potentials in the tests.)"""
    if dim == 2:
        return lambda *d: -0.5 / np.pi * d[axis] / sum(c * c for c in d)
    return lambda *d: -0.25 / np.pi * d[axis] / np.sqrt(sum(c * c for c in d)) ** 3


def laplace_source_derivative(dim, axis):
    """``AxisSourceDerivative(axis, LaplaceKernel(dim))``: ``-dG/dd_axis``."""
    g = laplace_target_derivative(dim, axis)
    return lambda *d: -g(*d)


def radial_power(dim, p):
    """|d|^p (no scaling constant): the term kernel the Helmholtz split of
    order > 1 tabulates in 3D (expansion_wrangler_fpnd.py:144-175)."""
    return lambda *d: sum(c * c for c in d) ** (0.5 * p)


def radial_power_log(dim, p):
    """|d|^p log|d|: the 2D term kernel of the same split (178-210)."""
    return lambda *d: sum(c * c for c in d) ** (0.5 * p) * 0.5 * np.log(sum(c * c for c in d))


def radial_power_target_derivative(dim, axis, p):
    """``d/dd_axis |d|^p = p |d|^(p-2) d_axis`` (the split term under a target
    derivative wrapper, expansion_wrangler_fpnd.py:5018-5035)."""
    return lambda *d: p * sum(c * c for c in d) ** (0.5 * p - 1.0) * d[axis]


def radial_power_log_target_derivative(dim, axis, p):
    """``d/dd_axis |d|^p log|d| = d_axis |d|^(p-2) (p log|d| + 1)`` (5037-5050)."""
    def g(*d):
        r2 = sum(c * c for c in d)
        return d[axis] * r2 ** (0.5 * p - 1.0) * (0.5 * p * np.log(r2) + 1.0)
    return g


def constant(dim):
    return lambda *d: np.ones_like(np.asarray(d[0], dtype=np.float64))


def helmholtz3d(k):
    def g(*d):
        r = np.sqrt(sum(c * c for c in d))
        return np.exp(1j * k * r) / (4 * np.pi * r)

    return g


def helmholtz2d(k):
    """(i/4) H0^(1)(k r), the sumpy 2D Helmholtz kernel (scipy's Hankel
    function: a library routine, so the kernel value itself needs no pin)."""
    from scipy.special import hankel1

    def g(*d):
        r = np.sqrt(sum(c * c for c in d))
        return 0.25j * hankel1(0, k * r)

    return g


def helmholtz_target_derivative(dim, axis, k):
    """``AxisTargetDerivative(axis, HelmholtzKernel(dim))``: ``G'(r) d_a / r``
    with ``G'`` = 2D ``-(ik/4) H1^(1)(kr)`` (scipy), 3D
    ``e^{ikr}(ikr - 1) / (4 pi r^2)`` -- the derivative of ``helmholtz2d`` /
    ``helmholtz3d``, pinned by central differences of those in the tests."""
    if dim == 2:
        from scipy.special import hankel1

        def g(*d):
            r = np.sqrt(sum(c * c for c in d))
            return -0.25j * k * hankel1(1, k * r) * d[axis] / r
    else:
        def g(*d):
            r = np.sqrt(sum(c * c for c in d))
            return np.exp(1j * k * r) * (1j * k * r - 1.0) / (4 * np.pi * r ** 3) * d[axis]

    return g


def helmholtz_source_derivative(dim, axis, k):
    """``AxisSourceDerivative(axis, HelmholtzKernel(dim))``: ``-dG/dd_axis``."""
    g = helmholtz_target_derivative(dim, axis, k)
    return lambda *d: -g(*d)


# ---------------------------------------------------------- quadratures
def box_quad_duffy_nd(func, bounds, singular_point, deg_regular, radial_order):
    """``singular_integral_2d.py:856-949`` (tanh-sinh-fast), vectorised."""
    dim = len(bounds)
    sp = np.array(
        [min(max(float(singular_point[i]), bounds[i][0]), bounds[i][1]) for i in range(dim)]
    )
    tail_x, tail_w = tail_nodes(dim - 1, deg_regular)
    rho, w_rho = tanh_sinh_fast(radial_order)
    total = 0.0
    for signs in itertools.product([-1.0, 1.0], repeat=dim):
        lengths = np.array(
            [sp[i] - bounds[i][0] if signs[i] < 0 else bounds[i][1] - sp[i] for i in range(dim)]
        )
        if np.any(lengths <= 0):
            continue
        scale = float(np.prod(lengths))
        for perm in itertools.permutations(range(dim)):
            for tx, tw in zip(tail_x, tail_w):
                rs = np.empty((dim, rho.size))
                rs[0] = rho
                if dim > 1:
                    rs[1:] = np.asarray(tx)[:, None]
                u = np.empty((dim, rho.size))
                cum = np.ones(rho.size)
                for i, ax in enumerate(perm):
                    cum = cum * rs[i]
                    u[ax] = cum
                x = sp[:, None] + np.asarray(signs)[:, None] * lengths[:, None] * u
                jac = scale * np.ones(rho.size)
                for i in range(dim - 1):
                    jac = jac * rs[i] ** (dim - 1 - i)
                vals = func(*x) * jac
                vals = np.where(np.isfinite(vals), vals, 0.0)
                total = total + tw * np.dot(w_rho, vals)
    return total


def tria_quad_duffy(func, tria, deg_theta, radial_order):
    """``singular_integral_2d.py:387-415, 594-691`` (tanh-sinh-fast): the
    triangle (p0, p1, p2) with the singularity at p0 is mapped affinely onto
    the template (0,0)-(1,0)-(0,1) and then by
    (a, b) = (rho cos^2 theta, rho sin^2 theta), Jacobian rho sin(2 theta),
    onto [0,1] x [0, pi/2]; GL in theta, the radial rule in rho."""
    p0, p1, p2 = (np.asarray(p, np.float64) for p in tria)
    area2 = abs((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]))
    if area2 == 0.0:
        return 0.0
    th, wth = np.polynomial.legendre.leggauss(int(deg_theta))
    th = 0.25 * np.pi * (th + 1.0)
    wth = 0.25 * np.pi * wth
    rho, wr = tanh_sinh_fast(radial_order)
    total = 0.0
    for theta, wt in zip(th, wth):
        a = rho * np.cos(theta) ** 2
        b = rho * np.sin(theta) ** 2
        x = p0[:, None] + (p1 - p0)[:, None] * a + (p2 - p0)[:, None] * b
        jac = rho * np.sin(2.0 * theta) * area2
        vals = func(x[0], x[1]) * jac
        vals = np.where(np.isfinite(vals), vals, 0.0)
        total += wt * np.dot(wr, vals)
    return total


def box_quad_duffy_2d_triangles(func, a, b, c, d, sp, deg_theta, radial_order):
    """``singular_integral_2d.py:731-761``: four triangles at the clamped point."""
    sp = (min(max(sp[0], a), b), min(max(sp[1], c), d))
    box = ((a, c), (b, c), (b, d), (a, d))
    total = 0.0
    for i in range(4):
        total += tria_quad_duffy(func, (sp, box[i], box[(i + 1) % 4]), deg_theta, radial_order)
    return total


# --------------------------------------------------------- table entries
class TableGeometry:
    """Geometry of a reference table on ``[0, H]^d`` (no values)."""

    def __init__(self, dim, q, H=1.0):
        from . import orbit_ref

        self.dim, self.q, self.H = int(dim), int(q), float(H)
        self.n = self.q**self.dim
        x, _ = np.polynomial.legendre.leggauss(self.q)
        self.xi = 0.5 * self.H * (x + 1.0)
        self.bw = bary_weights(self.xi)
        self.vecs, self.case_indices = orbit_ref.gallery(self.dim)
        grids = np.meshgrid(*([self.xi] * self.dim), indexing="ij")
        self.q_points = np.stack([g.ravel() for g in grids], axis=-1)

    def target_point(self, t, case_id):
        """``find_target_point`` (917-946)."""
        vec = np.asarray(self.vecs[case_id], np.float64)
        cntr = np.full(self.dim, 0.5 * self.H) + vec / 4.0 * self.H
        m = int(np.max(np.abs(vec)))
        size = 1.0 if m == 0 else (m - 2) / 2.0
        return cntr + size * (self.q_points[t] - 0.5 * self.H)

    def mode(self, s):
        from .orbit_ref import decode_axes

        idx = decode_axes(s, self.q, self.dim)

        def f(*coords):
            val = 1.0
            for k, c in enumerate(coords):
                val = val * lagrange_1d(self.xi, idx[k], c, self.bw)
            return val

        return f

    def entry(self, entry_id, kernel, deg_regular=20, radial_order=61, triangles_2d=False):
        """``compute_table_entry_duffy_radial`` (1099-1177), target - source."""
        n = self.n
        case_id = entry_id // (n * n)
        s = (entry_id % (n * n)) // n
        t = entry_id % n
        x_t = self.target_point(t, case_id)
        mode = self.mode(s)

        def integrand(*y):
            return mode(*y) * kernel(*[x_t[i] - y[i] for i in range(self.dim)])

        bounds = [(0.0, self.H)] * self.dim
        if self.dim == 2 and triangles_2d:
            return box_quad_duffy_2d_triangles(
                integrand, 0.0, self.H, 0.0, self.H, x_t, deg_regular, radial_order
            )
        return box_quad_duffy_nd(integrand, bounds, x_t, deg_regular, radial_order)


# ----------------------------------------------------------- closed forms
def box_potential_gradient(x, lo, hi, axis, h=1e-3):
    """``d/dx_axis`` of the closed-form uniform-box Laplace potential
    (``box_potential_laplace2d/3d``), by the 4th-order central difference
    with step ``h * max(hi - lo)``; the potential is smooth away from the
    box surface, so for a point at distance >> h from it the error is
    O(h^4) + rounding O(eps / h), about 1e-12 relative."""
    dim = len(x)
    f = box_potential_laplace2d if dim == 2 else box_potential_laplace3d
    step = h * float(np.max(np.asarray(hi) - np.asarray(lo)))

    def at(k):
        y = np.array(x, dtype=np.float64)
        y[axis] += k * step
        return f(y, lo, hi)

    return (8.0 * (at(1) - at(-1)) - (at(2) - at(-2))) / (12.0 * step)



def box_potential_laplace3d(x, lo, hi):
    """int_box 1/(4 pi |x - y|) dy for the box [lo, hi] (closed form).

    With X = y - x, the antiderivative of 1/r over a box is
    F = YZ ln(X+r) + XZ ln(Y+r) + XY ln(Z+r)
        - X^2/2 atan(YZ/(Xr)) - Y^2/2 atan(XZ/(Yr)) - Z^2/2 atan(XY/(Zr)),
    summed over the 8 corners with alternating signs.
    """
    total = 0.0
    for cx, cy, cz in itertools.product((0, 1), repeat=3):
        X = (hi[0] if cx else lo[0]) - x[0]
        Y = (hi[1] if cy else lo[1]) - x[1]
        Z = (hi[2] if cz else lo[2]) - x[2]
        r = np.sqrt(X * X + Y * Y + Z * Z)

        def ln(a):
            return np.log(a + r) if a + r > 0 else 0.0

        def at(num, den, coef):
            return 0.0 if coef == 0.0 or den == 0.0 else coef * np.arctan(num / den)

        F = Y * Z * ln(X) + X * Z * ln(Y) + X * Y * ln(Z)
        F -= 0.5 * at(Y * Z, X * r, X * X) + 0.5 * at(X * Z, Y * r, Y * Y) + 0.5 * at(
            X * Y, Z * r, Z * Z
        )
        sign = (-1) ** (3 - (cx + cy + cz))
        total += sign * F
    return total / (4 * np.pi)


def box_potential_laplace2d(x, lo, hi):
    """int_rect -log|x - y| / (2 pi) dy (closed form).

    With X = y - x: int int ln(X^2+Y^2) dX dY
      = XY ln(X^2+Y^2) - 3XY + X^2 atan(Y/X) + Y^2 atan(X/Y).
    """
    total = 0.0
    for cx, cy in itertools.product((0, 1), repeat=2):
        X = (hi[0] if cx else lo[0]) - x[0]
        Y = (hi[1] if cy else lo[1]) - x[1]
        r2 = X * X + Y * Y
        F = -3.0 * X * Y
        if r2 > 0:
            F += X * Y * np.log(r2)
        if X != 0:
            F += X * X * np.arctan(Y / X)
        if Y != 0:
            F += Y * Y * np.arctan(X / Y)
        total += (-1) ** (2 - (cx + cy)) * F
    return -0.5 * total / (2 * np.pi)
