"""mbirjax_amd: the mbirjax projector + VCD hot path on AMD Instinct MI355X.

Drop-in for mbirjax's TomographyModel API (ParallelBeamModel, ConeBeamModel, TranslationModel,
forward_project, back_project, sparse_forward_project, sparse_back_project,
compute_hessian_diagonal, recon, prox_map, set_params, get_params), computed by
hand-written gfx950 HIP kernels in libmbx.so (C ABI: include/mbx.h).
"""
from ._utils import Param
from .parameter_handler import ParameterHandler
from .tomography_model import TomographyModel, ReconParams
from .parallel_beam import ParallelBeamModel
from .cone_beam import ConeBeamModel
from .translation_model import TranslationModel
from .denoising import QGGMRFDenoiser, median_filter3d
from .qggmrf import (get_b_from_nbr_wts, qggmrf_gradient_and_hessian_at_indices, qggmrf_loss,
                     prox_gradient_at_indices)
from .vcd_utils import (get_2d_ror_mask, gen_pixel_partition, gen_set_of_pixel_partitions, gen_partition_sequence,
                        gen_full_indices)
from .tomography_utils import generate_direct_recon_filter
from . import utilities
from . import autograd

__all__ = ['Param', 'ParameterHandler', 'TomographyModel', 'ReconParams', 'ParallelBeamModel', 'ConeBeamModel',
           'TranslationModel', 'QGGMRFDenoiser', 'median_filter3d',
           'get_b_from_nbr_wts', 'qggmrf_gradient_and_hessian_at_indices', 'qggmrf_loss',
           'prox_gradient_at_indices', 'get_2d_ror_mask', 'gen_pixel_partition', 'gen_set_of_pixel_partitions',
           'gen_partition_sequence', 'gen_full_indices', 'generate_direct_recon_filter', 'utilities', 'autograd']
