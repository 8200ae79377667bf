"""SPH kernel types accepted by ``sphMapping`` / ``healpix_map``.

Mirrors the SPHKernels.jl types the reference dispatches on
(``kernel::AbstractSPHKernel``, cic_interpolation.jl:40; SPHKernels.jl compat
"2", an external dependency).  Only the type and ``dim`` travel to the
device: the shape functions are evaluated inline by the HIP kernels, and the
normalisation ``norm * h_inv^dim`` cancels in every deposit.
"""
from __future__ import annotations


class AbstractSPHKernel:
    kernel_id: int = -1
    name = "abstract"

    def __init__(self, dim: int = 3):
        if dim not in (1, 2, 3):
            raise ValueError("dim must be 1, 2 or 3")
        self.dim = int(dim)

    def __repr__(self):
        return f"{type(self).__name__}({self.dim})"


class Cubic(AbstractSPHKernel):
    kernel_id = 0


class Quintic(AbstractSPHKernel):
    kernel_id = 1


class WendlandC2(AbstractSPHKernel):
    kernel_id = 2


class WendlandC4(AbstractSPHKernel):
    kernel_id = 3


class WendlandC6(AbstractSPHKernel):
    kernel_id = 4


class WendlandC8(AbstractSPHKernel):
    kernel_id = 5


def kernel_args(kernel) -> tuple[int, int]:
    """(kernel id, kernel_dim) for the C ABI; kernel_dim 1 maps to 2 (it cancels)."""
    if not isinstance(kernel, AbstractSPHKernel):
        raise TypeError(f"kernel must be an AbstractSPHKernel, got {type(kernel).__name__}")
    return kernel.kernel_id, (kernel.dim if kernel.dim in (2, 3) else 2)
