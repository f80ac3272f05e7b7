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


def supports_inferred_scaling(kernel) -> bool:
    """``list1.py:263-276``: Laplace and constant kernels in 2D/3D."""
    return isinstance(kernel, (LaplaceKernel, ConstantKernel)) and kernel.dim in (2, 3)


def inferred_scaling(kernel):
    """``(power, log_displacement)`` of the exact single-table scaling.

    ``list1.py:278-355``: ``G(h r) = (h/H)^p G(H r)`` with ``p = 2`` for the
    Laplace volume potential in 2D and 3D (the 2D one adds the
    ``-(1/2pi) (h/H)^2 log(h/H) * mode_normalizer`` term) and ``p = dim`` for
    the constant kernel.
    """
    if isinstance(kernel, LaplaceKernel) and kernel.dim in (2, 3):
        return 2, kernel.dim == 2
    if isinstance(kernel, ConstantKernel) and kernel.dim in (2, 3):
        return kernel.dim, False
    raise RuntimeError(f"no inferred scaling rule for {kernel!r}")
