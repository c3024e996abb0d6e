"""MI355X (gfx950) laminography reconstruction hot path.

The compute runs in lamino_amd/liblamino_amd.so (hand-written HIP kernels
behind the C ABI of include/lamino.h); this package is the Python face of it.
"""
from . import _lib  # noqa: F401
from .api import *  # noqa: F401,F403

__all__ = [
    "Context", "DeviceArray", "PolarGrid", "beam_dir_vector", "forward", "adjoint", "nufft3d1", "nufft3d2",
    "sysmat", "gradient", "residual", "ToeplitzOperator", "fft2", "RampPreconditioner", "grad_u", "divergence",
    "laplacian", "qggmrf", "dot", "cgsolver", "split_bregman", "nagopt", "lipschitz", "ReconParams", "MBIR1",
    "MBIR2", "MBIR3", "autotune_bins", "mbir_footprint", "demag", "forward_multi", "adjoint_multi",
]
