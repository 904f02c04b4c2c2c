"""TranslationModel: a cone-beam source and flat detector with the object
translated, not rotated, between views (mbirjax/translation_model.py).

The projectors are the fan-beam kernels of libmbx (include/mbx.h,
MBX_GEOM_TRANSLATION): the per-(pixel, view) geometry table is built from the
view's translation vector instead of an angle, with the recon rows (y) at their
own pitch delta_recon_row.  This class supplies the geometry record, the psf
radius and the automatic recon size with the reference's float32 arithmetic.
"""
import math
import warnings
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from .tomography_model import TomographyModel


def _f32(x):
    return np.float32(x)


class TranslationModel(TomographyModel):
    """Translation tomography model (mbirjax/translation_model.py:12-67).

    Args:
        sinogram_shape (tuple): (views, rows, channels)
        translation_vectors (array): (views, 3) translations (x, y, z) in ALU;
            positive x shifts the object left, z up, y away from the source
        source_detector_dist (float): source to detector distance (ALU, finite)
        source_iso_dist (float): source to iso-plane distance (ALU)

    The row (y) direction is regularised less, qggmrf_nbr_wts = (0.5, 1, 1)
    (mbirjax/translation_model.py:63-66).
    """

    GEOM_KIND = _lib.GEOM_TRANSLATION

    def __init__(self, sinogram_shape, translation_vectors, source_detector_dist, source_iso_dist):
        self.entries_per_cylinder_batch = 128
        tv = translation_vectors.detach().cpu() if isinstance(translation_vectors, torch.Tensor) else translation_vectors
        tv = np.asarray(tv, dtype=np.float32)  # the reference holds them as a float32 array (jnp.asarray)
        super().__init__(sinogram_shape, translation_vectors=tv, source_detector_dist=source_detector_dist,
                         source_iso_dist=source_iso_dist, view_params_name='translation_vectors',
                         qggmrf_nbr_wts=(0.5, 1, 1))
        # The reference sets use_ror_mask = False before TomographyModel.__init__,
        # which then sets it to True (mbirjax/tomography_model.py:66); the
        # effective value, kept here, is True.

    def get_magnification(self):
        """source_detector_dist / source_iso_dist (mbirjax/translation_model.py:69-81)."""
        sdd, sid = self.get_params(['source_detector_dist', 'source_iso_dist'])
        if math.isinf(sdd):
            raise ValueError('Distance from source to detector is infinite, which means all translated projections '
                             'have the same information.')
        return sdd / sid

    def verify_valid_params(self):
        super().verify_valid_params()
        sinogram_shape, tv = self.get_params(['sinogram_shape', 'translation_vectors'])
        tv = np.asarray(tv)
        if tv.ndim != 2 or tv.shape[1] != 3 or tv.shape[0] != sinogram_shape[0]:
            raise ValueError("translation_vectors must have shape (num_views, 3); got {} for {} views".format(
                tv.shape, sinogram_shape[0]))

    def get_geometry_parameters(self):
        """namedtuple of the translation geometry (mbirjax/translation_model.py:83-107)."""
        names = ['delta_det_row', 'delta_det_channel', 'det_row_offset', 'det_channel_offset',
                 'source_detector_dist', 'delta_voxel', 'delta_recon_row']
        values = self.get_params(names)
        names += ['magnification', 'psf_radius', 'entries_per_cylinder_batch']
        values.append(self.get_magnification())
        values.append(self.get_psf_radius())
        values.append(self.entries_per_cylinder_batch)
        return namedtuple('GeometryParams', names)(*values)

    def get_psf_radius(self):
        """Tap radius from the largest magnification, that of the voxel
        closest to the source (mbirjax/translation_model.py:109-140)."""
        delta_det_row, delta_det_channel, sid, sdd, recon_shape, delta_voxel, delta_recon_row, tv = self.get_params(
            ['delta_det_row', 'delta_det_channel', 'source_iso_dist', 'source_detector_dist', 'recon_shape',
             'delta_voxel', 'delta_recon_row', 'translation_vectors'])
        self.get_magnification()  # raises for an infinite source_detector_dist
        delta_det = min(delta_det_row, delta_det_channel)
        max_translation = np.asarray(tv, dtype=np.float32).max(axis=0)
        closest = _f32(sid - (0.5 * recon_shape[0] * delta_recon_row) - max_translation[1])
        if closest < 0:
            raise ValueError('Reconstruction volume extends into source - no valid projection in this case.')
        max_mag = _f32(sdd / closest)
        psf_radius = int(np.ceil(_f32(np.ceil(_f32(delta_voxel * max_mag / delta_det)) / 2)))
        if psf_radius > 4:
            warnings.warn('A single voxel may project onto 100 or more detector elements, which may lead to '
                          'artifacts. Consider using smaller voxels.')
        return psf_radius

    def auto_set_recon_size(self, sinogram_shape, no_compile=True, no_warning=False):
        """Recon shape and row pitch from the detector, the cone angle and the
        translation extent (mbirjax/translation_model.py:142-186), float32."""
        sdd, sid = self.get_params(['source_detector_dist', 'source_iso_dist'])
        delta_det_row, delta_det_channel = self.get_params(['delta_det_row', 'delta_det_channel'])
        mag = self.get_magnification()
        delta_voxel = self.get_params('delta_voxel')
        num_views, num_det_rows, num_det_channels = sinogram_shape
        tv = np.asarray(self.get_params('translation_vectors'), dtype=np.float32)
        detect_box = np.array([delta_det_channel * num_det_channels, delta_det_row * num_det_rows], dtype=np.float32)
        half_cone_angle = np.arctan2(detect_box.max(), _f32(2 * sdd))
        delta_recon_row = float(np.sqrt(2) * (delta_voxel / np.tan(half_cone_angle)))
        cube = tv.max(axis=0) - tv.min(axis=0)
        recon_box = np.ceil((np.array([cube[0], cube[2]], dtype=np.float32) + detect_box / _f32(mag) / 2) /
                            _f32(delta_voxel))
        num_recon_rows = np.ceil(_f32(num_views * num_det_channels * num_det_rows) / (recon_box[0] * recon_box[1]))
        max_recon_rows = np.floor(_f32((sid - cube[1]) / delta_recon_row))
        if max_recon_rows < 1:
            print("[Error] Computed max_recon_rows = {} < 1. This suggests the object extends beyond the "
                  "source.".format(max_recon_rows))
        num_recon_rows = min(num_recon_rows, max_recon_rows)
        recon_shape = (int(num_recon_rows), int(recon_box[0]), int(recon_box[1]))
        self.set_params(no_compile=no_compile, no_warning=no_warning, recon_shape=recon_shape,
                        delta_recon_row=delta_recon_row)
