"""stablebear_amd: MI355X-native (gfx950) pairwise PCF integration.

Drop-in for stablebear's pairwise path: ``pdist`` (L_p distance matrices),
``l2_kernel`` (L2 Gram matrices), ``cdist`` (cross distances) and ``lp_norm``
(per-PCF norms) over tensors of piecewise constant functions, returning
``DistanceMatrix`` / ``SymmetricMatrix`` / ``FloatTensor`` in the reference's
layouts. All
integrals run in hand-written HIP kernels (``libstablebear_amd.so``, C-ABI in
``include/stablebear_amd.h``); there is no CPU compute path.

Extensions beyond the reference: ``precision="fp32c"`` (Kahan-compensated
FP32 with exact FP64 breakpoint handling), ``pdist_batched`` (many problems in
one launch, hipGraph-capturable), autograd (``stablebear_amd.autograd``) and
an autotuner for pair-tile shapes.
"""

from . import io, rows, system
from .io import load, save
from .batched import l2_kernel_batched, pdist_batched
from .distance import Distance_f32_f32, Distance_f64_f64, cdist, lp_distance, pdist
from .distance_matrix import DistanceMatrix
from .inner_product import InnerProduct_f32_f32, InnerProduct_f64_f64, l2_kernel
from .norms import Norms_f32_f32, Norms_f64_f64, lp_norm
from .pcf import Pcf, Rectangle, iterate_rectangles
from .symmetric_matrix import SymmetricMatrix
from .tensor import FloatTensor, IntPcfTensor, PcfContainerLike, PcfTensor, zeros
from .typing import dtype, float32, float64, int32, int64, pcf32, pcf32i, pcf64, pcf64i

__version__ = "0.1.0"

__all__ = [
    "Pcf", "PcfTensor", "IntPcfTensor", "FloatTensor", "PcfContainerLike", "zeros",
    "DistanceMatrix", "SymmetricMatrix", "Rectangle",
    "pdist", "cdist", "l2_kernel", "lp_distance", "lp_norm", "iterate_rectangles",
    "pdist_batched", "l2_kernel_batched",
    "Distance_f32_f32", "Distance_f64_f64", "InnerProduct_f32_f32", "InnerProduct_f64_f64",
    "Norms_f32_f32", "Norms_f64_f64",
    "dtype", "float32", "float64", "int32", "int64", "pcf32", "pcf64", "pcf32i", "pcf64i",
    "system", "io", "rows", "save", "load",
]
