"""QGGMRFDenoiser and median_filter3d (mbirjax/denoising.py).

The denoiser is VCD with the identity as forward model: every subset update
(qGGMRF gradient and Hessian, the step, the line search and both image
updates) runs in libmbx (mbx_denoise_subset_update, include/mbx.h) on the
device, one launch chain per subset with no host synchronisation; the host
drives the sweep over the partition and the stopping rule.
"""
import datetime
import warnings

import numpy as np
import torch

from . import _lib
from . import vcd_utils
from .qggmrf import get_b_from_nbr_wts
from .tomography_model import TomographyModel, ReconParams


class QGGMRFDenoiser(TomographyModel):
    """qGGMRF proximal-map denoiser of a 3D image (mbirjax/denoising.py:14-364).
    The primary interface is :meth:`denoise`."""

    GEOM_KIND = None

    def __init__(self, image_shape):
        super().__init__(image_shape, view_params_name='None', sigma_noise=None)
        self.use_ror_mask = False
        self.set_params(sharpness=0)  # the denoiser's default sharpness
        self.set_params(granularity=[16], partition_sequence=[0])  # one fixed partition of 16 subsets
        self.set_params(use_gpu='full')

    def get_magnification(self):
        return 1.0

    def verify_valid_params(self):
        super().verify_valid_params()
        sinogram_shape, image_shape = self.get_params(['sinogram_shape', 'recon_shape'])
        if tuple(image_shape) != tuple(sinogram_shape):
            raise ValueError("image_shape and sinogram_shape must be the same. \n"
                             "Got {} for image_shape and {} for sinogram_shape".format(image_shape, sinogram_shape))

    def get_geometry_parameters(self):
        raise NotImplementedError('get_geometry_parameters is not implemented for QGGMRFDenoiser.')

    def create_projectors(self):
        pass

    def auto_set_recon_size(self, sinogram_shape, no_compile=True, no_warning=False):
        self.set_params(no_compile=no_compile, no_warning=no_warning, recon_shape=tuple(sinogram_shape))

    def auto_set_sigma_y(self, sinogram, sino_indicator, weights=1):
        self.set_params(no_warning=True, sigma_y=self.get_params('sigma_noise'), auto_regularize_flag=True)

    def estimate_image_noise_std(self, image):
        """Noise standard deviation of the image from local differences
        (mbirjax/denoising.py:80-91)."""
        image = _host(image)
        support = self._get_sino_indicator(image, sigma_noise=0.0)
        sigma_noise = self._get_estimate_of_recon_std(image, support)
        support = self._get_sino_indicator(image, sigma_noise=sigma_noise)
        return self._get_estimate_of_recon_std(image, support)

    def _get_estimate_of_recon_std(self, noisy_image, support_indicator):
        """Mean over the support of the std of each voxel and its three lower
        neighbours (mbirjax/denoising.py:93-116)."""
        inds = np.where(support_indicator)
        inds = [np.clip(inds[j], 1, None) for j in range(len(inds))]
        vals = np.stack([noisy_image[inds[0], inds[1], inds[2]], noisy_image[inds[0] - 1, inds[1], inds[2]],
                         noisy_image[inds[0], inds[1] - 1, inds[2]], noisy_image[inds[0], inds[1], inds[2] - 1]],
                        axis=0)
        return np.mean(np.std(vals, axis=0))

    def _get_sino_indicator(self, noisy_image, sigma_noise=None):
        """Support of the image by thresholding (mbirjax/denoising.py:118-137)."""
        if sigma_noise is None:
            sigma_noise = self.get_params('sigma_noise')
        threshold = (0.01 * 5.0) * np.mean(np.fabs(noisy_image)) + sigma_noise
        threshold = min(threshold, np.amax(noisy_image))
        return np.int8(noisy_image >= threshold)

    def recon(self, *args, **kwargs):
        raise NotImplementedError('recon is not implemented for QGGMRFDenoiser.  Use `denoise` instead.')

    def denoise(self, image, sigma_noise=None, use_ror_mask=False, init_image=None, max_iterations=15,
                stop_threshold_change_pct=0.2, first_iteration=0, logfile_path='./logs/recon.log', print_logs=True):
        """MAP estimate of the image under the qGGMRF prior with AWGN of
        standard deviation sigma_noise (estimated from the image if None),
        by VCD sweeps over one fixed partition (mbirjax/denoising.py:142-364).

        Returns (denoised_image (device tensor), denoiser_dict)."""
        _lib.require_gpu()
        self.use_ror_mask = use_ror_mask
        host_image = _host(image)
        if sigma_noise is None:
            sigma_noise = self.estimate_image_noise_std(host_image)
        self.set_params(sigma_noise=sigma_noise)
        if first_iteration == 0 or self.logger is None:
            self.setup_logger(logfile_path=logfile_path, print_logs=print_logs)
        self._log('info', 'Initializing QGGMRFDenoiser')
        regularization_params = self.auto_set_regularization_params(host_image)

        image_shape, granularity, partition_sequence = self.get_params(['recon_shape', 'granularity',
                                                                        'partition_sequence'])
        rows, cols, nz = image_shape
        partition = vcd_utils.gen_set_of_pixel_partitions(image_shape, [granularity[partition_sequence[0]]],
                                                          use_ror_mask=self.use_ror_mask)[0]
        dev = self._device()
        img = self._tensor(image).reshape(rows * cols, nz)
        x = (img.clone() if init_image is None else self._tensor(init_image).reshape(rows * cols, nz).clone())
        err = (img - x).contiguous()  # identity forward model: error image = image - estimate
        subsets = [torch.as_tensor(np.asarray(s, np.int32), device=dev) for s in partition]

        fm_constant = 1.0 / float(self.get_params('sigma_y')) ** 2
        nbr, sigma_x, p, q, T = self.get_params(['qggmrf_nbr_wts', 'sigma_x', 'p', 'q', 'T'])
        qp = _lib.qggmrf_struct(get_b_from_nbr_wts(nbr), sigma_x, p, q, T)
        ws_bytes = _lib.lib().mbx_denoise_workspace_bytes(max(int(s.numel()) for s in subsets), nz)
        ws = torch.empty(int(ws_bytes), dtype=torch.uint8, device=dev)
        stats = torch.zeros(4, dtype=torch.float64, device=dev)
        L, st = _lib.lib(), _lib.stream()

        self._log('info', 'Starting VCD iterations')
        stop = stop_threshold_change_pct / 100.0
        nmae_update, alpha_values = [], []
        num_iters = 0
        for it in range(max_iterations):
            stats.zero_()
            for s in subsets:
                _lib.check(L.mbx_denoise_subset_update(_lib.ptr(x), _lib.ptr(err), rows, cols, nz, _lib.ptr(s),
                                                       int(s.numel()), fm_constant, qp, _lib.ptr(stats),
                                                       _lib.ptr(ws), int(ws_bytes), st))
            ell1, alpha_sum, count = stats[:3].tolist()
            nmae = ell1 / float(torch.sum(torch.abs(x.double())).item())
            nmae_update.append(nmae)
            alpha_values.append(alpha_sum / max(count, 1.0))
            num_iters = it + 1
            if it % 5 == 0:
                self._log('info', 'After iteration {} of a max of {}: Pct change={:.4f}'.format(
                    it + first_iteration, max_iterations, 100 * nmae))
            if nmae < stop:
                break

        recon_param_values = [int(num_iters), granularity, partition_sequence, None, None, regularization_params,
                              [100 * float(v) for v in nmae_update], [float(v) for v in alpha_values]]
        recon_params = ReconParams(*tuple(recon_param_values))._asdict()
        notes = 'Reconstruction completed: {}\n\n'.format(datetime.datetime.now())
        denoiser_dict = self.get_recon_dict(recon_params, notes=notes)
        return x.reshape(rows, cols, nz), denoiser_dict


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def median_filter3d(x, max_block_gb=4.0):
    """27-point (3x3x3) median filter with edge-replicated borders
    (mbirjax/denoising.py:414-496), computed by mbx_median_filter3d.  Float64
    input stays float64; every other dtype is filtered in float32.  Returns a
    device tensor of x's shape.  max_block_gb is accepted for the reference's
    signature; the kernel needs no blocking."""
    _lib.require_gpu()
    dev = torch.device('cuda', torch.cuda.current_device())
    t = x.detach() if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    if t.dim() != 3:
        raise ValueError('median_filter3d expects a 3D array')
    dtype = torch.float64 if t.dtype == torch.float64 else torch.float32
    src = t.to(device=dev, dtype=dtype).contiguous()
    out = torch.empty_like(src)
    code = _lib.DTYPE_F64 if dtype == torch.float64 else _lib.DTYPE_F32
    d0, d1, d2 = src.shape
    _lib.check(_lib.lib().mbx_median_filter3d(_lib.ptr(src), code, d0, d1, d2, _lib.ptr(out), _lib.stream()))
    return out
