"""GPU parity of the qGGMRF denoiser and the 3x3x3 median filter.

Bars: the median filter equals the oracle (and the reference's golden 3x3x3
example) exactly; the denoiser after a fixed number of sweeps is within
NRMSE 1e-3 of the FP64 oracle run on the same seeded partition; on the
reference's own case (tests/test_denoiser.py:31-53) the denoised Shepp-Logan
phantom is within nrmse 0.2 of the clean one.
"""
import numpy as np
import pytest
import torch

import mbirjax_amd as mj
from oracle import denoise as oden

pytestmark = pytest.mark.gpu


def test_median_golden_gpu(gpu):
    c = np.array([[[3., 3., 4.], [6., 6., 7.], [6., 7., 8.]],
                  [[10., 11., 11.], [12., 13., 14.], [15., 15., 16.]],
                  [[18., 19., 20.], [19., 20., 20.], [22., 23., 23.]]])
    b = mj.median_filter3d(np.arange(27).reshape((3, 3, 3)))
    assert np.array_equal(b.cpu().numpy(), c)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_median_vs_oracle(gpu, dtype):
    x = np.random.default_rng(20).standard_normal((13, 9, 17)).astype(dtype)
    out = mj.median_filter3d(torch.as_tensor(x, device='cuda'))
    assert out.dtype == (torch.float64 if dtype == np.float64 else torch.float32)
    assert np.array_equal(out.cpu().numpy(), oden.median_filter3d(x).astype(dtype))
    # thin volumes: every voxel on a border
    y = np.random.default_rng(21).random((1, 2, 5)).astype(dtype)
    assert np.array_equal(mj.median_filter3d(y).cpu().numpy(), oden.median_filter3d(y).astype(dtype))


def test_denoiser_vs_oracle(gpu):
    """16 x 16 x 12 noisy ellipsoid image (256 pixels: 16 subsets of 16, no
    padding), sigma_noise 0.05, sharpness 0, 10 sweeps, against the oracle with
    the same partition and its own automatic sigma_x."""
    shape = (16, 16, 12)
    clean = mj.utilities.gen_random_ellipsoid_phantom(shape, seed=3)
    noisy = (clean + 0.05 * np.random.default_rng(4).standard_normal(shape)).astype(np.float32)
    d = mj.QGGMRFDenoiser(shape)
    np.random.seed(0)
    out, info = d.denoise(noisy, sigma_noise=0.05, max_iterations=10, stop_threshold_change_pct=0.0,
                          print_logs=False, logfile_path=None)
    np.random.seed(0)
    ref, _ = oden.denoise(noisy, 0.05, max_iterations=10, stop_threshold_change_pct=0.0)
    assert abs(info['recon_params']['regularization_params']['sigma_x'] - oden.sigma_x_auto(noisy, 0.05, 0.0)) <= \
        1e-6 * oden.sigma_x_auto(noisy, 0.05, 0.0)
    r = out.cpu().numpy().astype(np.float64)
    nrmse = np.linalg.norm(r - ref) / np.linalg.norm(ref)
    print("denoiser NRMSE vs oracle %.3e; vs clean %.3f (noisy %.3f)" % (
        nrmse, np.linalg.norm(r - clean) / np.linalg.norm(clean),
        np.linalg.norm(noisy - clean) / np.linalg.norm(clean)), flush=True)
    assert nrmse <= 1e-3


def test_denoiser_reference_case(gpu):
    """mbirjax tests/test_denoiser.py:31-53: 100^3 low-dynamic-range Shepp-Logan
    phantom + N(0, 0.1^2), sigma_noise 0.1, sharpness 0, 20 iterations,
    stop at 0.1 %: nrmse against the clean phantom below 0.2."""
    np.random.seed(0)
    shape = (100, 100, 100)
    phantom = np.asarray(mj.utilities.generate_3d_shepp_logan_low_dynamic_range(shape))
    noisy = phantom + 0.1 * np.random.randn(*shape)
    d = mj.QGGMRFDenoiser(phantom.shape)
    d.set_params(sharpness=0.0)
    out, _ = d.denoise(noisy, 0.1, max_iterations=20, stop_threshold_change_pct=0.1, print_logs=False,
                       logfile_path=None)
    nrmse = np.linalg.norm(out.cpu().numpy() - phantom) / np.linalg.norm(phantom)
    print("reference denoiser case nrmse %.4f" % nrmse, flush=True)
    assert nrmse < 0.2
