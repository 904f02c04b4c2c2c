"""ParallelBeamModel: parallel-beam geometry (mbirjax/parallel_beam.py).

Slice k maps to detector row k; the projector arithmetic runs in libmbx.
"""
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from .tomography_model import TomographyModel


class ParallelBeamModel(TomographyModel):
    """Parallel-beam model (mbirjax/parallel_beam.py:13-52).

    Args:
        sinogram_shape (tuple): (views, rows, channels)
        angles (array): view angles in radians, one per view
    """

    GEOM_KIND = _lib.GEOM_PARALLEL

    def __init__(self, sinogram_shape, angles):
        angles = np.asarray(angles.detach().cpu() if isinstance(angles, torch.Tensor) else angles)
        super().__init__(sinogram_shape, angles=angles, view_params_name='angles')

    def get_magnification(self):
        return 1.0

    def verify_valid_params(self):
        super().verify_valid_params()
        sinogram_shape, angles = self.get_params(['sinogram_shape', 'angles'])
        if np.shape(angles)[0] != sinogram_shape[0]:
            raise ValueError("Number view dependent parameter vectors must equal the number of views. \n"
                             "Got {} for length of view-dependent parameters and {} for number of views.".format(
                                 np.shape(angles)[0], sinogram_shape[0]))
        recon_shape = self.get_params('recon_shape')
        if recon_shape[2] != sinogram_shape[1]:
            raise ValueError("Number of recon slices must match number of sinogram rows. \n"
                             "Got {} for recon_shape and {} for sinogram_shape".format(recon_shape, sinogram_shape))

    def get_geometry_parameters(self):
        """(mbirjax/parallel_beam.py:94-113)"""
        names = ['delta_det_channel', 'det_channel_offset', 'delta_voxel']
        values = self.get_params(names)
        names += ['psf_radius']
        values.append(self.get_psf_radius())
        return namedtuple('GeometryParams', names)(*values)

    def get_psf_radius(self):
        delta_det_channel, delta_voxel = self.get_params(['delta_det_channel', 'delta_voxel'])
        return int(np.ceil(np.ceil(np.float32(delta_voxel / delta_det_channel)) / np.float32(2)))

    def auto_set_recon_size(self, sinogram_shape, no_compile=True, no_warning=False):
        """(mbirjax/parallel_beam.py:125-136) with float32 rounding."""
        delta_det_row, delta_det_channel, delta_voxel = self.get_params(['delta_det_row', 'delta_det_channel',
                                                                         'delta_voxel'])
        R, C = sinogram_shape[1:3]
        mag = self.get_magnification()
        rows = int(np.ceil(np.float32(C * delta_det_channel / (delta_voxel * mag))))
        slices = int(np.round(np.float32(R * ((delta_det_row / delta_voxel) / mag))))
        self.set_params(no_compile=no_compile, no_warning=no_warning, recon_shape=(rows, rows, slices))

    def direct_recon(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.fbp_recon(sinogram, filter_name=filter_name, view_batch_size=view_batch_size)

    def direct_filter(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.fbp_filter(sinogram, filter_name=filter_name, view_batch_size=view_batch_size)

    def fbp_filter(self, sinogram, filter_name="ramp", view_batch_size=100):
        """Ramp filtering of each detector row (mbirjax/parallel_beam.py:303-348)."""
        from .tomography_utils import generate_direct_recon_filter
        s = self._tensor(sinogram)
        V, R, C = s.shape
        if view_batch_size is None:
            view_batch_size = 128
        h = generate_direct_recon_filter(C, filter_name=filter_name, device=s.device)
        h = h * (1.0 / self.get_params('delta_voxel') ** 2)
        out = torch.empty_like(s)
        for i in range(0, V, view_batch_size):
            out[i:i + view_batch_size] = self._ramp_conv_rows(s[i:i + view_batch_size], h)
        return out * (np.pi / V)

    def fbp_recon(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.back_project(self.fbp_filter(sinogram, filter_name=filter_name, view_batch_size=view_batch_size))
