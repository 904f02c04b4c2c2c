"""ConeBeamModel: cone-beam geometry with a flat detector (mbirjax/cone_beam.py).

The projector arithmetic runs in libmbx (include/mbx.h); this class supplies
the geometry record, the psf radii and automatic recon size with the
reference's float32 rounding, and the FDK filter.
"""
import math
import warnings
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from .tomography_model import TomographyModel


def _f32_ceil(x):
    return float(np.ceil(np.float32(x)))


class ConeBeamModel(TomographyModel):
    """Cone-beam model (mbirjax/cone_beam.py:16-60).

    Args:
        sinogram_shape (tuple): (views, rows, channels)
        angles (array): view angles in radians, one per view
        source_detector_dist (float): source to detector distance (ALU)
        source_iso_dist (float): source to rotation axis distance (ALU)
    """

    GEOM_KIND = _lib.GEOM_CONE

    def __init__(self, sinogram_shape, angles, source_detector_dist, source_iso_dist):
        self.bp_psf_radius = 1
        self.entries_per_cylinder_batch = 128
        self.slice_range_length = 0
        angles = np.asarray(angles.detach().cpu() if isinstance(angles, torch.Tensor) else angles)
        super().__init__(sinogram_shape, angles=angles, source_detector_dist=source_detector_dist,
                         source_iso_dist=source_iso_dist, view_params_name='angles', recon_slice_offset=0.0)

    def get_magnification(self):
        sdd, sid = self.get_params(['source_detector_dist', 'source_iso_dist'])
        if math.isinf(sdd):
            return 1
        return sdd / sid

    def verify_valid_params(self):
        super().verify_valid_params()
        sinogram_shape, angles = self.get_params(['sinogram_shape', 'angles'])
        if np.shape(angles)[0] != sinogram_shape[0]:
            raise ValueError("Number view dependent parameter vectors must equal the number of views. \n"
                             "Got {} for length of view-dependent parameters and {} for number of views.".format(
                                 np.shape(angles)[0], sinogram_shape[0]))
        sdd, delta_det_row, det_row_offset = self.get_params(['source_detector_dist', 'delta_det_row',
                                                              'det_row_offset'])
        if delta_det_row * sinogram_shape[1] / 2 + abs(det_row_offset) > sdd:
            warnings.warn('Cone angle is more than 45 degrees.  This will likely produce recon artifacts.')

    def get_geometry_parameters(self):
        """namedtuple of the cone geometry (mbirjax/cone_beam.py:110-136)."""
        names = ['delta_det_row', 'delta_det_channel', 'det_row_offset', 'det_channel_offset',
                 'source_detector_dist', 'delta_voxel', 'recon_slice_offset']
        values = self.get_params(names)
        names += ['magnification', 'psf_radius', 'bp_psf_radius', 'entries_per_cylinder_batch',
                  'slice_range_length']
        values.append(self.get_magnification())
        values.append(self.get_psf_radius())
        values.append(self.bp_psf_radius)
        values.append(self.entries_per_cylinder_batch)
        values.append(self.slice_range_length)
        return namedtuple('GeometryParams', names)(*values)

    def get_psf_radius(self):
        """Tap radii with the reference's float32 ceil (mbirjax/cone_beam.py:138-172)."""
        delta_det_row, delta_det_channel, sdd, recon_shape, delta_voxel = self.get_params(
            ['delta_det_row', 'delta_det_channel', 'source_detector_dist', 'recon_shape', 'delta_voxel'])
        mag = self.get_magnification()
        delta_det = min(delta_det_row, delta_det_channel)
        if math.isinf(sdd):
            max_mag = min_mag = 1
        else:
            sid = sdd / mag
            extent = max(recon_shape[0], recon_shape[1]) * delta_voxel
            max_mag = sdd / (sid - extent)
            min_mag = sdd / (sid + extent)
        psf_radius = int(_f32_ceil(_f32_ceil(delta_voxel * max_mag / delta_det) / 2))
        max_voxels_per_detector = delta_det / (min_mag * delta_voxel)
        self.bp_psf_radius = int(_f32_ceil(_f32_ceil(max_voxels_per_detector) / 2))
        self.slice_range_length = int(1 + 2 * self.bp_psf_radius +
                                      _f32_ceil(self.entries_per_cylinder_batch * max_voxels_per_detector))
        return psf_radius

    def auto_set_recon_size(self, sinogram_shape, no_compile=True, no_warning=False):
        """(mbirjax/cone_beam.py:174-187) with float32 round-half-even."""
        delta_det_row, delta_det_channel, delta_voxel = self.get_params(['delta_det_row', 'delta_det_channel',
                                                                         'delta_voxel'])
        R, C = sinogram_shape[1:3]
        mag = self.get_magnification()
        rows = int(np.round(np.float32(C * ((delta_det_channel / delta_voxel) / mag))))
        slices = int(np.round(np.float32(R * ((delta_det_row / delta_voxel) / mag))))
        self.set_params(no_compile=no_compile, no_warning=no_warning, recon_shape=(rows, rows, slices))

    # ------------------------------------------------------------------ FDK
    def direct_recon(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.fdk_recon(sinogram, filter_name=filter_name, view_batch_size=view_batch_size)

    def direct_filter(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.fdk_filter(sinogram, filter_name=filter_name, view_batch_size=view_batch_size)

    def fdk_filter(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        """Cosine pre-weighting and ramp filtering of each detector row
        (mbirjax/cone_beam.py:779-846)."""
        from .tomography_utils import generate_direct_recon_filter
        s = self._tensor(sinogram)
        V, R, C = s.shape
        sdd = self.get_params('source_detector_dist')
        delta_voxel, delta_det_row, delta_det_channel = self.get_params(['delta_voxel', 'delta_det_row',
                                                                         'delta_det_channel'])
        det_row_offset, det_channel_offset = self.get_params(['det_row_offset', 'det_channel_offset'])
        dev = s.device
        m = torch.arange(R, device=dev, dtype=torch.float32)
        n = torch.arange(C, device=dev, dtype=torch.float32)
        v = (m - (R - 1) / 2.0) * delta_det_row - det_row_offset
        u = (n - (C - 1) / 2.0) * delta_det_channel - det_channel_offset
        weight = sdd / torch.sqrt(sdd ** 2 + u[None, :] ** 2 + v[:, None] ** 2)
        h = generate_direct_recon_filter(C, filter_name=filter_name, device=dev)
        h = (delta_det_row / (delta_voxel ** 3 * self.get_magnification())) * h
        out = torch.empty_like(s)
        for i in range(0, V, view_batch_size):
            out[i:i + view_batch_size] = self._ramp_conv_rows(s[i:i + view_batch_size] * weight[None], h)
        return out * (np.pi / V)

    def fdk_recon(self, sinogram, filter_name="ramp", view_batch_size=TomographyModel.DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.back_project(self.fdk_filter(sinogram, filter_name=filter_name, view_batch_size=view_batch_size))
