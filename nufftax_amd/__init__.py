"""nufftax_amd: nufftax's NUFFT API on MI355X (gfx950) HIP kernels.

Drop-in for the public API of nufftax (nufftax/__init__.py:55-99): the nine
transforms, the type-3 grid-size helpers and the centered-FFT utilities, on
PyTorch tensors; ``nufftax_amd.core`` mirrors ``nufftax.core`` (spread /
interp with their VJPs, deconvolution, kernel helpers).  The compute path is the nufftx HIP library (libnufftx.so,
C ABI in include/nufftx.h); there is no CPU fallback.
"""

__version__ = "0.1.0"

from .params import (  # noqa: E402
    KernelParams,
    compute_grid_size,
    compute_kernel_params,
    compute_type3_grid_size,
    compute_type3_grid_sizes_2d,
    compute_type3_grid_sizes_3d,
    next_smooth_int,
)
from .transforms import (  # noqa: E402
    nufft1d1,
    nufft1d1_jvp,
    nufft1d2,
    nufft1d2_jvp,
    nufft1d3,
    nufft1d3_jvp,
    nufft2d1,
    nufft2d1_jvp,
    nufft2d2,
    nufft2d2_jvp,
    nufft2d3,
    nufft2d3_jvp,
    nufft3d1,
    nufft3d1_jvp,
    nufft3d2,
    nufft3d2_jvp,
    nufft3d3,
    nufft3d3_jvp,
)
from .utils import fftc, from_complex, ifftc, to_complex  # noqa: E402
from . import core  # noqa: E402

__all__ = [
    "nufft1d1", "nufft2d1", "nufft3d1",
    "nufft1d2", "nufft2d2", "nufft3d2",
    "nufft1d3", "nufft2d3", "nufft3d3",
    "nufft1d1_jvp", "nufft1d2_jvp", "nufft1d3_jvp",
    "nufft2d1_jvp", "nufft2d2_jvp", "nufft2d3_jvp",
    "nufft3d1_jvp", "nufft3d2_jvp", "nufft3d3_jvp",
    "compute_type3_grid_size", "compute_type3_grid_sizes_2d", "compute_type3_grid_sizes_3d",
    "compute_kernel_params", "compute_grid_size", "next_smooth_int", "KernelParams",
    "fftc", "ifftc", "to_complex", "from_complex", "core",
]
