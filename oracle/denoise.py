"""ORACLE -- test infrastructure only (see oracle/__init__.py).

NumPy FP64 restatement of mbirjax/denoising.py:
  median_filter3d                           denoising.py:414-496
  QGGMRFDenoiser._get_sino_indicator        denoising.py:118-137
  QGGMRFDenoiser._get_estimate_of_recon_std denoising.py:93-116
  QGGMRFDenoiser.estimate_image_noise_std   denoising.py:80-91
  QGGMRFDenoiser.denoise                    denoising.py:142-364 (sigma_y = sigma_noise,
                                            sigma_x from TomographyModel.auto_set_regularization_params,
                                            tomography_model.py:756-796)
  vcd_subset_denoiser                       denoising.py:366-411

Pinned by (tests/test_oracle.py): the reference's golden 3x3x3 median
(tests/test_denoiser.py:55-67), equality with scipy.ndimage.median_filter
(size 3, mode 'nearest'), the quadratic-prior (p = q = 2) fixed point against
a direct solve of (I/sigma^2 + L) x = y/sigma^2, and a constant image as a
fixed point.  The step size (prior over-relaxation 1.1 and forward Hessian 1,
denoising.py:379-398) does not move the fixed point: it is pinned only by
this restatement and by the reference's accuracy bar (nrmse < 0.2,
tests/test_denoiser.py:31-53), which the GPU test applies.
"""
import numpy as np

from . import partitions, qggmrf

EPS32 = float(np.finfo(np.float32).eps)


def median_filter3d(x):
    """Median of the 27 edge-replicated neighbours of every voxel."""
    x = np.asarray(x)
    d0, d1, d2 = x.shape
    xp = np.pad(x, 1, mode='edge')
    stack = [xp[1 + a:1 + a + d0, 1 + b:1 + b + d1, 1 + c:1 + c + d2]
             for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    return np.median(np.stack(stack, axis=0), axis=0)


def sino_indicator(image, sigma_noise):
    threshold = (0.01 * 5.0) * np.mean(np.fabs(image)) + sigma_noise
    threshold = min(threshold, np.amax(image))
    return np.int8(image >= threshold)


def recon_std(image, indicator):
    inds = [np.clip(i, 1, None) for i in np.where(indicator)]
    vals = np.stack([image[inds[0], inds[1], inds[2]], image[inds[0] - 1, inds[1], inds[2]],
                     image[inds[0], inds[1] - 1, inds[2]], image[inds[0], inds[1], inds[2] - 1]], axis=0)
    return np.mean(np.std(vals, axis=0))


def estimate_image_noise_std(image):
    s = recon_std(image, sino_indicator(image, 0.0))
    return recon_std(image, sino_indicator(image, s))


def sigma_x_auto(image, sigma_noise, sharpness):
    """sigma_x of auto_set_regularization_params on <= 20 sub-sampled slabs along axis 0."""
    step = image.shape[0] // min(20, image.shape[0])
    small = np.asarray(image[::step], dtype=np.float32)
    std = recon_std(small, sino_indicator(small, sigma_noise))
    return float(np.float32(0.2 * (2 ** sharpness) * std))


def subset_update(flat_image, flat_error, pix, fm_constant, qparams, shape):
    """One vcd_subset_denoiser step (FP64); returns ell1, alpha and updates in place."""
    pg, ph = qggmrf.gradient_and_hessian_at_indices(flat_image, shape, pix, qparams)
    e = flat_error[pix]
    fg = -fm_constant * e
    delta = -((fg + pg) / (1.0 + ph))
    pl = np.sum(pg * delta)
    pq = (1 / 1.1) * np.sum(ph * delta ** 2)
    fl = fm_constant * np.sum(e * delta)
    fq = fm_constant * np.sum(delta * delta)
    alpha = float(np.clip((fl - pl) / (fq + pq + EPS32), EPS32, 1.5))
    np.add.at(flat_image, pix, alpha * delta)
    flat_error[pix] = e - alpha * delta
    return float(np.sum(np.abs(alpha * delta))), alpha


def denoise(image, sigma_noise, sharpness=0.0, p=2.0, q=1.2, T=1.0, nbr_wts=(1.0, 1.0, 1.0), max_iterations=15,
            stop_threshold_change_pct=0.2, use_ror_mask=False, sigma_x=None, init_image=None):
    """The denoise loop: one partition of 16 subsets (drawn from the global
    NumPy RNG), swept in order each iteration."""
    image = np.asarray(image, np.float64)
    shape = image.shape
    rows, cols, nz = shape
    if sigma_x is None:
        sigma_x = sigma_x_auto(image, sigma_noise, sharpness)
    b = qggmrf.get_b_from_nbr_wts(nbr_wts)
    qparams = (b, float(sigma_x), float(p), float(q), float(T))
    partition = partitions.gen_set_of_pixel_partitions(shape, [16], use_ror_mask=use_ror_mask)[0]
    x = (image.copy() if init_image is None else np.asarray(init_image, np.float64).copy()).reshape(-1, nz)
    err = image.reshape(-1, nz) - x
    fmc = 1.0 / float(np.float32(sigma_noise)) ** 2
    nmae_hist = []
    for it in range(max_iterations):
        ell1 = 0.0
        for subset in partition:
            l1, _ = subset_update(x, err, np.asarray(subset), fmc, qparams, shape)
            ell1 += l1
        nmae = ell1 / np.sum(np.abs(x))
        nmae_hist.append(nmae)
        if nmae < stop_threshold_change_pct / 100.0:
            break
    return x.reshape(shape), nmae_hist
