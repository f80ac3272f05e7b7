"""Integral kernel descriptors for the near-field path.

The reference names kernels with sumpy objects (``LaplaceKernel(dim)``,
``HelmholtzKernel(dim)``, the test-only ``ConstantKernel``).  This package
needs only what the list-1 path consumes: the dimension, whether values are
complex, the reference's ``repr`` (used as the per-kernel table key, see
``expansion_wrangler_fpnd.py:2368-2404``), the scaling rule of
``list1.py:278-355`` and the device kernel id of the table builder.

Kernel conventions (sumpy expression times global scaling constant):

* Laplace 2D ``-log|d| / (2 pi)``, 3D ``1 / (4 pi |d|)``
* Helmholtz 2D ``(i/4) H0^(1)(k|d|)``, 3D ``e^{ik|d|} / (4 pi |d|)``
* Constant ``1``
* Axis derivatives of Laplace: target ``dG/dd_a`` = 2D ``-d_a / (2 pi |d|^2)``,
  3D ``-d_a / (4 pi |d|^3)`` (``d = target - source``); source = minus that
* Axis derivatives of Helmholtz: target ``G'(|d|) d_a / |d|`` with
  ``G'`` = 2D ``-(ik/4) H1^(1)(k|d|)``, 3D ``e^{ik|d|}(ik|d| - 1) / (4 pi |d|^2)``;
  source = minus that
* Radial split terms: ``|d|^p`` and ``|d|^p log|d|``
"""

from __future__ import annotations

from dataclasses import dataclass

from . import _capi


@dataclass(frozen=True)
class LaplaceKernel:
    dim: int
    is_complex_valued = False
    kind = _capi.KERNEL_LAPLACE

    def __repr__(self):
        return f"LapKnl{self.dim}D"

    def get_base_kernel(self):
        return self


@dataclass(frozen=True)
class ConstantKernel:
    dim: int
    is_complex_valued = False
    kind = _capi.KERNEL_CONSTANT

    def __repr__(self):
        return f"CstKnl{self.dim}D"

    def get_base_kernel(self):
        return self


@dataclass(frozen=True)
class HelmholtzKernel:
    dim: int
    k: float = 1.0
    is_complex_valued = True
    kind = _capi.KERNEL_HELMHOLTZ

    def __repr__(self):
        return f"HelmKnl{self.dim}D"

    def get_base_kernel(self):
        return self


@dataclass(frozen=True)
class RadialPowerKernel:
    """``|d|^power`` (real, no scaling constant): a term that the Helmholtz
    split of order > 1 takes out of the online remainder into a table (3D
    odd powers; ``expansion_wrangler_fpnd.py:144-175``)."""

    dim: int
    power: int
    is_complex_valued = False
    kind = _capi.KERNEL_POWER

    def __post_init__(self):
        if not 0 <= int(self.power) <= 32:
            raise ValueError("power must be 0..32")

    def __repr__(self):
        return f"RadialPowerKernel(dim={self.dim}, power={self.power})"

    def get_base_kernel(self):
        return self


@dataclass(frozen=True)
class RadialPowerLogKernel:
    """``|d|^power log|d|`` (real, no scaling constant, power > 0): the 2D
    terms of the Helmholtz split of order > 1
    (``expansion_wrangler_fpnd.py:178-210``)."""

    dim: int
    power: int
    is_complex_valued = False
    kind = _capi.KERNEL_POWER_LOG

    def __post_init__(self):
        if not 1 <= int(self.power) <= 32:
            raise ValueError("power must be 1..32 for r**power * log(r)")

    def __repr__(self):
        return f"RadialPowerLogKernel(dim={self.dim}, power={self.power})"

    def get_base_kernel(self):
        return self


@dataclass(frozen=True)
class AxisTargetDerivative:
    """``d/dx_axis`` of ``inner_kernel`` at the target (sumpy's wrapper of the
    same name): ``G_a(d) = dG/dd_a`` with ``d = target - source``."""

    axis: int
    inner_kernel: object

    @property
    def kind(self):
        """Table-builder kernel id (one derivative of Laplace, Helmholtz or a
        radial split term)."""
        return _derivative_kind(self, _capi.KERNEL_LAPLACE_TARGET_DERIV,
                                _capi.KERNEL_HELMHOLTZ_TARGET_DERIV,
                                _capi.KERNEL_POWER_TARGET_DERIV,
                                _capi.KERNEL_POWER_LOG_TARGET_DERIV)

    @property
    def dim(self):
        return self.inner_kernel.dim

    @property
    def is_complex_valued(self):
        return self.inner_kernel.is_complex_valued

    def __repr__(self):
        return f"AxisTargetDerivative({self.axis}, {self.inner_kernel!r})"

    def get_base_kernel(self):
        return self.inner_kernel.get_base_kernel()


@dataclass(frozen=True)
class AxisSourceDerivative:
    """``d/dy_axis`` of ``inner_kernel`` at the source: ``-dG/dd_a``."""

    axis: int
    inner_kernel: object

    @property
    def kind(self):
        return _derivative_kind(self, _capi.KERNEL_LAPLACE_SOURCE_DERIV,
                                _capi.KERNEL_HELMHOLTZ_SOURCE_DERIV,
                                _capi.KERNEL_POWER_SOURCE_DERIV,
                                _capi.KERNEL_POWER_LOG_SOURCE_DERIV)

    @property
    def dim(self):
        return self.inner_kernel.dim

    @property
    def is_complex_valued(self):
        return self.inner_kernel.is_complex_valued

    def __repr__(self):
        return f"AxisSourceDerivative({self.axis}, {self.inner_kernel!r})"

    def get_base_kernel(self):
        return self.inner_kernel.get_base_kernel()


_DERIVATIVES = (AxisTargetDerivative, AxisSourceDerivative)


def _derivative_kind(kernel, laplace_kind, helmholtz_kind, power_kind, power_log_kind):
    inner = kernel.inner_kernel
    if isinstance(inner, LaplaceKernel):
        return laplace_kind
    if isinstance(inner, HelmholtzKernel):
        return helmholtz_kind
    if isinstance(inner, RadialPowerKernel):
        if int(inner.power) < 1:
            raise ValueError("the derivative of |d|^0 is zero: no table")
        return power_kind
    if isinstance(inner, RadialPowerLogKernel):
        return power_log_kind
    raise NotImplementedError(f"no table builder for {kernel!r} (nested derivatives)")


def derivative_axes(kernel):
    """``(fixed_axes, axis_sign_power)`` of a kernel's axis-derivative chain.

    ``orbit_arithmetic.py:176-197``: every axis derivative pins its axis (no
    axis permutation may move it) and toggles that axis's sign power (a flip
    of the axis negates the entry).  Scalar kernels give ``((), zeros)``.
    """
    dim = kernel.dim
    fixed = set()
    power = [0] * dim
    k = kernel
    while isinstance(k, _DERIVATIVES):
        if not 0 <= int(k.axis) < dim:
            raise ValueError(f"derivative axis {k.axis} out of range for dim {dim}")
        fixed.add(int(k.axis))
        power[int(k.axis)] ^= 1
        k = k.inner_kernel
    return tuple(sorted(fixed)), tuple(power)


def _single_laplace_derivative(kernel, dim=None):
    return (isinstance(kernel, _DERIVATIVES)
            and isinstance(kernel.inner_kernel, LaplaceKernel)
            and (dim is None or kernel.dim == dim))


def supports_inferred_scaling(kernel) -> bool:
    """``list1.py:263-276``: Laplace and constant kernels in 2D/3D, and one
    axis target/source derivative of Laplace."""
    if _single_laplace_derivative(kernel):
        return kernel.dim in (2, 3)
    return isinstance(kernel, (LaplaceKernel, ConstantKernel)) and kernel.dim in (2, 3)


def inferred_scaling(kernel):
    """``(power, log_displacement)`` of the exact single-table scaling.

    ``list1.py:278-355``: ``G(h r) = (h/H)^p G(H r)`` with ``p = 2`` for the
    Laplace volume potential in 2D and 3D (the 2D one adds the
    ``-(1/2pi) (h/H)^2 log(h/H) * mode_normalizer`` term), ``p = 1`` and no
    displacement for an axis derivative of Laplace (``grad G`` is
    homogeneous of degree ``1 - d``), and ``p = dim`` for the constant kernel.
    """
    if _single_laplace_derivative(kernel) and kernel.dim in (2, 3):
        return 1, False
    if isinstance(kernel, LaplaceKernel) and kernel.dim in (2, 3):
        return 2, kernel.dim == 2
    if isinstance(kernel, ConstantKernel) and kernel.dim in (2, 3):
        return kernel.dim, False
    raise RuntimeError(f"no inferred scaling rule for {kernel!r}")
