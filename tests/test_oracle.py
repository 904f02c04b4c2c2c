"""Pins the CPU oracle (no GPU): golden vectors, closed forms and invariants.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) fails it.
"""
import json
import os

import numpy as np
import pytest

from oracle import denoise as oden
from oracle import partitions as opart
from oracle import projectors as op
from oracle import qggmrf as oq

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "qggmrf_golden.json")


# ---------------------------------------------------------------- qGGMRF
def load_golden():
    with open(GOLDEN) as f:
        g = json.load(f)
    params = (tuple(g["b"]), g["sigma_x"], g["p"], g["q"], g["T"])
    return g, params


def test_qggmrf_golden_vectors():
    """Reference golden grad/hess (mbirjax tests/test_qggmrf.py:83-134), allclose at rtol 1e-5."""
    g, params = load_golden()
    x = np.array(g["flat_recon"])
    grad, hess = oq.gradient_and_hessian_at_indices(x, g["recon_shape"], np.arange(9), params)
    assert np.allclose(grad, np.array(g["grad_ref"]), rtol=1e-5, atol=1e-6)
    assert np.allclose(hess, np.array(g["hess_ref"]), rtol=1e-5, atol=1e-6)


def test_qggmrf_golden_catches_wrong_neighbour():
    """A swapped neighbour order (col+1 weight on row+1) must be visible with unequal b."""
    g, params = load_golden()
    x = np.array(g["flat_recon"])
    b = (0.3, 0.1, 0.2, 0.15, 0.15, 0.1)
    p2 = (b,) + params[1:]
    g1, _ = oq.gradient_and_hessian_at_indices(x, g["recon_shape"], np.arange(9), p2)
    b_sw = (0.2, 0.15, 0.3, 0.1, 0.15, 0.1)
    g2, _ = oq.gradient_and_hessian_at_indices(x, g["recon_shape"], np.arange(9), (b_sw,) + params[1:])
    assert not np.allclose(g1, g2)


def test_b_tilde_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(5):
        p = rng.random() + 1
        q = p - 0.9 * rng.random()
        T = 0.5 + rng.random()
        sx = 0.1 + rng.random()
        d = 0.1 + rng.random(10)
        a = oq.get_2_b_tilde(d, 1.0, ((1,) * 6, sx, p, q, T)) / 2
        # rho'(d)/(2d) for rho(d) = |d|^p/(p sx^p) * s^(q-p)/(1+s^(q-p)), s = |d|/(T sx)
        h = 1e-6
        rho = lambda t: np.abs(t) ** p / (p * sx ** p) * (np.abs(t) / (T * sx)) ** (q - p) / (
            1 + (np.abs(t) / (T * sx)) ** (q - p))
        deriv = (rho(d + h) - rho(d - h)) / (2 * h)
        assert np.allclose(a, deriv / (2 * d), rtol=1e-6)
        assert np.allclose(a, oq.b_tilde_by_definition(d, sx, p, q, T), rtol=1e-5)


def test_qggmrf_gradient_is_loss_derivative():
    """FP64 finite differences of the prior loss (mbirjax tests/test_qggmrf.py:136-172)."""
    rng = np.random.default_rng(2)
    params = (tuple(oq.get_b_from_nbr_wts([1.0, 1.0, 1.0])), 0.789, 2.0413, 1.124, 1.46)
    shape = (4, 5, 6)
    x0 = rng.random(shape)
    d = rng.random(shape)
    grad, hess = oq.gradient_and_hessian_at_indices(x0.reshape(-1, shape[2]), shape, np.arange(20), params)
    eps = 1e-7
    fd = (oq.qggmrf_loss(x0 + eps * d, params) - oq.qggmrf_loss(x0 - eps * d, params)) / (2 * eps)
    assert np.allclose(fd, np.sum(grad * d.reshape(-1, shape[2])), rtol=1e-6)
    gf, hf = oq.full_grad_and_hessian(x0, params)
    assert np.allclose(gf.reshape(-1, shape[2]), grad, rtol=1e-6, atol=1e-9)  # eps32 shift vs clip
    assert np.allclose(hf.reshape(-1, shape[2]), hess, rtol=1e-6, atol=1e-9)


# ---------------------------------------------------------------- projectors
def cone_geom(V=12, R=8, C=12, sdd=48.0, sid=24.0, **kw):
    return op.Geometry("cone", (V, R, C), np.linspace(0, 2 * np.pi, V, endpoint=False), sdd, sid, **kw)


def par_geom(V=12, R=6, C=10, **kw):
    return op.Geometry("parallel", (V, R, C), np.linspace(0, np.pi, V, endpoint=False), **kw)


def test_psf_radii_at_benchmark_shapes():
    """psf radii are 1 and 1 at the cone benchmark geometries; recon sizes as documented."""
    g2 = op.Geometry("cone", (128, 128, 192), np.zeros(128), 768, 384)
    assert g2.recon_shape == (192, 192, 128)
    assert g2.psf_radii() == (1, 1)
    g4 = op.Geometry("cone", (720, 512, 768), np.zeros(720), 3072, 1536, recon_shape=(512, 512, 512))
    assert g4.psf_radii() == (1, 1)
    g1 = op.Geometry("parallel", (64, 64, 64), np.zeros(64))
    assert g1.recon_shape == (64, 64, 64) and g1.psf_radii() == (1, 1)


def trans_geom(V=6, R=8, C=12, sdd=48.0, sid=24.0, recon_shape=(4, 10, 8), delta_recon_row=1.3, seed=0, **kw):
    rng = np.random.default_rng(seed)
    tv = np.stack([rng.uniform(-2, 2, V), rng.uniform(-1, 1, V), rng.uniform(-1.5, 1.5, V)], axis=1)
    return op.Geometry("translation", (V, R, C), tv, sdd, sid, recon_shape=recon_shape,
                       delta_recon_row=delta_recon_row, **kw)


@pytest.mark.parametrize("geom", [cone_geom(), cone_geom(det_row_offset=0.7, det_channel_offset=-1.3,
                                                         recon_slice_offset=0.4), par_geom(),
                                  trans_geom(), trans_geom(det_row_offset=0.6, det_channel_offset=-0.9, seed=1)])
def test_adjoint_fp64(geom):
    rng = np.random.default_rng(3)
    rows, cols, nz = geom.recon_shape
    pix = np.arange(rows * cols)
    x = rng.standard_normal((pix.size, nz))
    y = rng.standard_normal(geom.sinogram_shape)
    lhs = np.sum(y * op.forward(geom, x, pix))
    rhs = np.sum(op.back(geom, y, pix) * x)
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + 1)


@pytest.mark.parametrize("geom", [cone_geom(V=5, R=5, C=6, sdd=24.0, sid=12.0), par_geom(V=5, R=3, C=5),
                                  trans_geom(V=3, R=5, C=6, sdd=24.0, sid=12.0, recon_shape=(2, 5, 4))])
def test_hessian_is_squared_column_norm(geom):
    A = op.dense_matrix(geom)
    rows, cols, nz = geom.recon_shape
    h = op.back(geom, np.ones(geom.sinogram_shape), np.arange(rows * cols), coeff_power=2)
    assert np.allclose(h.reshape(-1), np.sum(A * A, axis=0), rtol=1e-12, atol=1e-14)
    # and the back projector is the transpose of the dense forward matrix
    y = np.random.default_rng(4).standard_normal(A.shape[0])
    bt = op.back(geom, y.reshape(geom.sinogram_shape), np.arange(rows * cols)).reshape(-1)
    assert np.allclose(bt, A.T @ y, rtol=1e-12, atol=1e-12)


def test_parallel_closed_form_axis_sums():
    """At 0 and 90 degrees with delta_voxel == delta_det_channel and cols == C
    the parallel projector is an exact axis sum times delta_voxel."""
    C, R = 8, 4
    geom = op.Geometry("parallel", (2, R, C), np.array([0.0, np.pi / 2]))
    rows, cols, nz = geom.recon_shape
    assert (rows, cols, nz) == (C, C, R)
    x = np.random.default_rng(5).random((rows, cols, nz))
    s = op.forward(geom, x.reshape(-1, nz), np.arange(rows * cols))
    # view 0: channel n = column j, sum over rows i
    assert np.allclose(s[0], x.sum(axis=0).T * geom.delta_voxel, rtol=1e-12)
    # view 90 deg: channel n = C-1-i, sum over columns j (the float32 angle of
    # pi/2 is off by 4.4e-8 rad, which moves the taps by ~1e-7)
    assert np.allclose(s[1], x.sum(axis=1)[::-1].T * geom.delta_voxel, rtol=1e-6)


def test_cone_far_source_equals_parallel():
    V, R, C = 7, 6, 10
    ang = np.linspace(0, np.pi, V, endpoint=False)
    gc = op.Geometry("cone", (V, R, C), ang, 1e9, 1e9)
    gp = op.Geometry("parallel", (V, R, C), ang)
    assert gc.recon_shape == gp.recon_shape
    rows, cols, nz = gp.recon_shape
    x = np.random.default_rng(6).random((rows * cols, nz))
    pix = np.arange(rows * cols)
    assert np.allclose(op.forward(gc, x, pix), op.forward(gp, x, pix), rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_translation_shift_equivalence(axis):
    """Translating the object by one voxel pitch along x (cols), y (rows, at
    delta_recon_row) or z (slices) equals shifting the voxel array by one:
    pins the sign, axis and pitch of each translation component."""
    # translations and pitches exact in float32 (the view parameters are
    # held as float32), so that t + pitch is exact
    rng = np.random.default_rng(2)
    tv = rng.integers(-64, 64, size=(4, 3)) / 32.0
    geom = op.Geometry("translation", (4, 8, 12), tv, 48.0, 24.0, recon_shape=(4, 9, 7), delta_recon_row=1.25)
    rows, cols, nz = geom.recon_shape
    x = np.random.default_rng(8).random((rows, cols, nz))
    arr_axis, pitch = {0: (1, geom.delta_voxel), 1: (0, geom.delta_recon_row), 2: (2, geom.delta_voxel)}[axis]
    idx = [slice(None)] * 3
    idx[arr_axis] = -1
    x[tuple(idx)] = 0.0
    xs = np.roll(x, 1, axis=arr_axis)  # voxel j -> j + 1 along the axis
    shifted = op.Geometry("translation", geom.sinogram_shape, geom.angles + pitch * np.eye(3)[axis],
                          geom.source_detector_dist, geom.source_iso_dist, recon_shape=geom.recon_shape,
                          delta_recon_row=geom.delta_recon_row)
    pix = np.arange(rows * cols)
    a = op.forward(geom, x.reshape(-1, nz), pix)
    b = op.forward(shifted, xs.reshape(-1, nz), pix)
    assert np.max(np.abs(a)) > 0
    assert np.allclose(a, b, rtol=1e-9, atol=1e-12)


def test_translation_far_source_is_parallel_over_dv():
    """With the source at infinity (magnification 1) and no translation, the
    translation projector is the parallel projector at angle 0 without its
    delta_voxel factor (mbirjax/translation_model.py:276-284 vs
    parallel_beam.py:223-263)."""
    V, R, C, dv = 1, 5, 8, 2.0
    common = dict(recon_shape=(3, C, R), delta_voxel=dv, delta_det_channel=dv, delta_det_row=dv)
    gt = op.Geometry("translation", (V, R, C), np.zeros((V, 3)), 1e9, 1e9, delta_recon_row=1.0, **common)
    gp = op.Geometry("parallel", (V, R, C), np.zeros(V), **common)
    rows, cols, nz = gt.recon_shape
    x = np.random.default_rng(9).random((rows * cols, nz))
    pix = np.arange(rows * cols)
    assert np.allclose(op.forward(gt, x, pix), op.forward(gp, x, pix) / dv, rtol=1e-6, atol=1e-9)


def test_view_subsets_are_additive():
    geom = cone_geom()
    rows, cols, nz = geom.recon_shape
    pix = np.arange(rows * cols)
    rng = np.random.default_rng(7)
    x = rng.random((pix.size, nz))
    full = op.forward(geom, x, pix)
    views = np.array([3, 0, 7, 7, 11])
    sub = op.forward(geom, x, pix, views=views)
    assert np.array_equal(sub, full[views])
    y = rng.random(geom.sinogram_shape)
    parts = [np.arange(j, 12, 3) for j in range(3)]
    total = sum(op.back(geom, y, pix, views=v) for v in parts)
    assert np.allclose(total, op.back(geom, y, pix), rtol=1e-12)


# ---------------------------------------------------------------- partitions
def test_partitions_structure_and_counts():
    np.random.seed(0)
    shape = (64, 64, 8)
    mask = opart.get_2d_ror_mask(shape)
    P = int(mask.sum())
    parts = opart.gen_set_of_pixel_partitions(shape, [1, 2, 4, 8, 16, 32, 64, 128, 256])
    for K, part in zip([1, 2, 4, 8, 16, 32, 64, 128, 256], parts):
        assert part.shape == (K, int(np.ceil(P / K)))
        assert np.all(np.diff(part, axis=1) >= 0)  # each subset sorted
        uniq, counts = np.unique(part, return_counts=True)
        assert uniq.size == P and np.all(mask.ravel()[uniq])
        pad = part.size - P
        assert np.sum(counts - 1) == pad  # the padding repeats pad distinct pixels once each
        assert np.all(counts <= 2)
    # ROR pixel counts at the benchmark shapes
    assert int(opart.get_2d_ror_mask((192, 192)).sum()) == 28600
    assert int(opart.get_2d_ror_mask((512, 512)).sum()) == 205012
    assert int(opart.get_2d_ror_mask((64, 64)).sum()) == 3096


# ---------------------------------------------------------------- denoiser
def test_median_golden_vector():
    """The reference's 3x3x3 example (mbirjax tests/test_denoiser.py:55-67)."""
    c = np.array([[[3., 3., 4.], [6., 6., 7.], [6., 7., 8.]],
                  [[10., 11., 11.], [12., 13., 14.], [15., 15., 16.]],
                  [[18., 19., 20.], [19., 20., 20.], [22., 23., 23.]]])
    assert np.array_equal(oden.median_filter3d(np.arange(27).reshape(3, 3, 3)), c)


def test_median_equals_scipy_nearest():
    from scipy import ndimage
    x = np.random.default_rng(10).standard_normal((9, 7, 5))
    assert np.array_equal(oden.median_filter3d(x), ndimage.median_filter(x, size=3, mode='nearest'))


def test_denoiser_quadratic_fixed_point():
    """With p = q = 2, 2 b~(delta) = b / (2 sigma_x^2) for every delta, so the
    prior gradient is L x with L the graph Laplacian of the 6-neighbourhood with
    edge weights b_axis / (2 sigma_x^2) (clipped in-plane borders and reflected
    slice ends contribute no edge), and the denoiser's limit solves
    (I/sigma^2 + L) x = y/sigma^2.  rows*cols is a multiple of the 16 subsets
    (no padded duplicates).  The step's absolute eps32 in the denominator
    slows the last digits, so the bar is 3e-5 after 50 sweeps; a wrong
    Laplacian weight or sign is off by O(1)."""
    from scipy.sparse import lil_matrix
    shape = (8, 4, 3)
    rows, cols, nz = shape
    rng = np.random.default_rng(11)
    y = rng.random(shape)
    sigma, sigma_x = 0.3, 0.25
    b = oq.get_b_from_nbr_wts((1.0, 0.7, 1.3))
    n = y.size
    idx = np.arange(n).reshape(shape)
    Lap = lil_matrix((n, n))
    for axis, wb in ((0, b[0]), (1, b[2]), (2, b[4])):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis], hi[axis] = slice(0, -1), slice(1, None)
        for u, v in zip(idx[tuple(lo)].ravel(), idx[tuple(hi)].ravel()):
            w = wb / (2 * sigma_x ** 2)
            Lap[u, u] += w
            Lap[v, v] += w
            Lap[u, v] -= w
            Lap[v, u] -= w
    sigma = float(np.float32(sigma))
    A = (np.eye(n) / sigma ** 2) + Lap.toarray()
    x_star = np.linalg.solve(A, y.ravel() / sigma ** 2).reshape(shape)
    np.random.seed(0)
    x, _ = oden.denoise(y, np.float32(sigma), p=2.0, q=2.0, nbr_wts=(1.0, 0.7, 1.3), sigma_x=sigma_x,
                        max_iterations=50, stop_threshold_change_pct=0.0)
    assert np.linalg.norm(x - x_star) <= 3e-5 * np.linalg.norm(x_star)


def test_denoiser_constant_image_is_fixed():
    np.random.seed(1)
    y = np.full((6, 5, 4), 0.7)
    x, nmae = oden.denoise(y, 0.1, sigma_x=0.05, max_iterations=3, stop_threshold_change_pct=0.0)
    assert np.allclose(x, y, rtol=0, atol=1e-12)
