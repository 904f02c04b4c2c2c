"""ORACLE -- test infrastructure only (see oracle/__init__.py).

FP64 restatement of the direct-reconstruction filters:
  generate_direct_recon_filter  mbirjax/tomography_utils.py:4-35
  ConeBeamModel.fdk_filter      mbirjax/cone_beam.py:779-846
  ParallelBeamModel.fbp_filter  mbirjax/parallel_beam.py:303-348
The 'valid' fftconvolve of each detector row with the (2C-1)-tap ramp is
restated as an exact Toeplitz product out[i] = sum_j row[j] h(i - j).
Pinned by: the ramp closed form (h(0) = 1/4, h(odd) = -1/(pi n)^2, h(even != 0)
= 0) and FBP/FDK reconstruction tolerances of tests/test_fbp_fdk.py:46-47.
"""
import numpy as np


def ramp_filter(num_channels):
    n = np.arange(-num_channels + 1, num_channels, dtype=np.float64)
    return 0.5 * np.sinc(n) - 0.25 * np.sinc(n / 2) ** 2


def _toeplitz_rows(sino, h):
    C = sino.shape[-1]
    i = np.arange(C)
    H = h[(i[:, None] - i[None, :]) + (C - 1)]  # H[i, j] = h(i - j)
    return np.asarray(sino, np.float64) @ H.T


def fbp_filter(geom, sinogram):
    V, R, C = sinogram.shape
    h = ramp_filter(C) / geom.delta_voxel ** 2
    return _toeplitz_rows(sinogram, h) * (np.pi / V)


def fdk_filter(geom, sinogram):
    V, R, C = sinogram.shape
    sdd = geom.source_detector_dist
    m = np.arange(R)
    n = np.arange(C)
    mg, ng = np.meshgrid(m, n, indexing="ij")
    v = (mg - (R - 1) / 2.0) * geom.delta_det_row - geom.det_row_offset
    u = (ng - (C - 1) / 2.0) * geom.delta_det_channel - geom.det_channel_offset
    weight = sdd / np.sqrt(sdd ** 2 + u ** 2 + v ** 2)
    alpha = geom.delta_det_row / (geom.delta_voxel ** 3 * geom.magnification())
    h = alpha * ramp_filter(C)
    return _toeplitz_rows(np.asarray(sinogram, np.float64) * weight[None], h) * (np.pi / V)
