"""GPU parity of the qGGMRF kernel and of VCD reconstruction against the oracle.

Bars: qGGMRF golden vectors at the reference's allclose (rtol 1e-5); VCD
reconstructions after a fixed iteration count within NRMSE <= 1e-3 of the FP64
oracle run with the same seeded partitions.
"""
import json
import os

import numpy as np
import pytest
import torch

import mbirjax_amd as mj
from oracle import partitions as opart
from oracle import projectors as op
from oracle import qggmrf as oq
from oracle import vcd as ovcd

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "qggmrf_golden.json")


def test_qggmrf_golden_gpu(gpu):
    with open(GOLDEN) as f:
        g = json.load(f)
    params = (tuple(g["b"]), g["sigma_x"], g["p"], g["q"], g["T"])
    grad, hess = mj.qggmrf_gradient_and_hessian_at_indices(np.array(g["flat_recon"], np.float32), (3, 3, 3),
                                                           np.arange(9), params)
    assert np.allclose(grad.cpu().numpy(), np.array(g["grad_ref"]), rtol=1e-5, atol=1e-6)
    assert np.allclose(hess.cpu().numpy(), np.array(g["hess_ref"]), rtol=1e-5, atol=1e-6)


def test_qggmrf_random_vs_oracle(gpu):
    rng = np.random.default_rng(21)
    shape = (17, 13, 70)
    x = rng.random((shape[0] * shape[1], shape[2])).astype(np.float32)
    params = (mj.get_b_from_nbr_wts([1.0, 2.0, 0.5]), 0.3, 2.0, 1.2, 1.0)
    idx = rng.choice(shape[0] * shape[1], size=60, replace=False)
    g, h = mj.qggmrf_gradient_and_hessian_at_indices(x, shape, idx, params)
    go, ho = oq.gradient_and_hessian_at_indices(x.astype(np.float64), shape, idx, params)
    assert np.linalg.norm(g.cpu().numpy() - go) / np.linalg.norm(go) <= 1e-5
    assert np.linalg.norm(h.cpu().numpy() - ho) / np.linalg.norm(ho) <= 1e-5


def problem(kind):
    if kind == "cone":
        V, R, C = 24, 16, 32
        ang = np.linspace(0, 2 * np.pi, V, endpoint=False)
        m = mj.ConeBeamModel((V, R, C), ang, source_detector_dist=4 * C, source_iso_dist=2 * C)
        g = op.Geometry("cone", (V, R, C), ang, 4 * C, 2 * C)
    else:
        V, R, C = 32, 16, 32
        ang = np.linspace(0, np.pi, V, endpoint=False)
        m = mj.ParallelBeamModel((V, R, C), ang)
        g = op.Geometry("parallel", (V, R, C), ang)
    assert m.get_params('recon_shape') == g.recon_shape
    phantom = mj.utilities.gen_random_ellipsoid_phantom(g.recon_shape, seed=0)
    rows, cols, nz = g.recon_shape
    ror = np.flatnonzero(opart.get_2d_ror_mask(g.recon_shape))
    sino = op.forward(g, phantom.reshape(-1, nz)[ror], ror)
    noise = np.random.default_rng(1).standard_normal(sino.shape)
    sino = (sino + 0.01 * np.max(sino) * noise).astype(np.float32)
    return m, g, phantom, sino


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_vcd_recon_matches_oracle(gpu, kind):
    print("vcd parity", kind, flush=True)
    m, g, phantom, sino = problem(kind)
    m.set_params(sharpness=0.0, verbose=0)
    np.random.seed(5)
    recon, rd = m.recon(sino, max_iterations=6, stop_threshold_change_pct=0, print_logs=False, logfile_path=None)
    recon = recon.cpu().numpy().astype(np.float64)
    om = ovcd.OracleModel(g, sharpness=0.0)
    np.random.seed(5)
    ref, stats = om.recon(sino, max_iterations=6, stop_threshold_change_pct=0)
    nrmse = np.linalg.norm(recon - ref) / np.linalg.norm(ref)
    print("%s recon NRMSE vs oracle %.3e, vs phantom %.3f" % (
        kind, nrmse, np.linalg.norm(recon - phantom) / np.linalg.norm(phantom)), flush=True)
    assert nrmse <= 1e-3
    rp = rd['recon_params']
    assert rp['num_iterations'] == 6
    assert np.allclose(rp['fm_rmse'], stats['fm_rmse'], rtol=1e-3)
    assert np.allclose(rp['alpha_values'], stats['alpha_values'], rtol=1e-3, atol=1e-6)
    assert np.isclose(rp['regularization_params']['sigma_y'], stats['regularization_params']['sigma_y'], rtol=1e-7)


def test_error_sinogram_invariant(gpu):
    """e == y - A x after the reconstruction (the overcomplete VCD state)."""
    m, g, phantom, sino = problem("cone")
    m.set_params(sharpness=0.0, verbose=0)
    np.random.seed(2)
    recon, _ = m.recon(sino, max_iterations=4, stop_threshold_change_pct=0, print_logs=False, logfile_path=None)
    e = m.from_channel_major(m.last_error_sinogram.float())
    Ax = m.forward_project(recon)
    y = torch.as_tensor(sino, device='cuda')
    resid = y - Ax
    assert (torch.linalg.norm(resid - e) / torch.linalg.norm(y)).item() <= 1e-5


def test_recon_quality_reference_tolerances(gpu):
    """Phantom tolerances of mbirjax tests/test_vcd.py:21-22 at its geometry
    (64 views x 40 rows x 128 channels; cone with sdd = sid = 4*128)."""
    V, R, C = 64, 40, 128
    for kind, tol in (("parallel", (0.15, 0.38, 0.04)), ("cone", (0.19, 0.56, 0.05))):
        if kind == "cone":
            cone_angle = 2 * np.arctan2(C / 2, 4 * C)
            ang = np.linspace(-(np.pi + cone_angle) / 2, (np.pi + cone_angle) / 2, V, endpoint=False)
            m = mj.ConeBeamModel((V, R, C), ang, source_detector_dist=4 * C, source_iso_dist=4 * C)
        else:
            ang = np.linspace(-np.pi / 2, np.pi / 2, V, endpoint=False)
            m = mj.ParallelBeamModel((V, R, C), ang)
        np.random.seed(0)
        phantom = mj.utilities.generate_3d_shepp_logan_low_dynamic_range(m.get_params('recon_shape'))
        sino = m.forward_project(phantom)
        m.set_params(verbose=0)
        recon, _ = m.recon(sino, print_logs=False, logfile_path=None)
        r = recon.cpu().numpy()
        nrmse = np.linalg.norm(r - phantom) / np.linalg.norm(phantom)
        max_diff = np.max(np.abs(r - phantom))
        p95 = np.percentile(np.abs(r - phantom), 95)
        print(kind, "nrmse %.3f max %.3f p95 %.3f" % (nrmse, max_diff, p95), flush=True)
        assert nrmse < tol[0] and max_diff < tol[1] and p95 < tol[2]


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_weighted_positivity_recon_matches_oracle(gpu, kind):
    """Transmission weights and the positivity constraint (the weighted and
    clamped forms of the subset update, mbirjax/tomography_model.py:1398-1401,
    1497-1507) against the FP64 oracle."""
    m, g, phantom, sino = problem(kind)
    w = np.exp(-sino / np.max(sino)).astype(np.float32)
    assert np.allclose(m.gen_weights(sino / np.max(sino), 'transmission').cpu().numpy(), w, rtol=1e-6)
    m.set_params(sharpness=0.0, verbose=0, positivity_flag=True)
    np.random.seed(8)
    recon, rd = m.recon(sino, weights=w, max_iterations=5, stop_threshold_change_pct=0, print_logs=False,
                        logfile_path=None)
    recon = recon.cpu().numpy().astype(np.float64)
    om = ovcd.OracleModel(g, sharpness=0.0, positivity_flag=True)
    np.random.seed(8)
    ref, stats = om.recon(sino, weights=w, max_iterations=5, stop_threshold_change_pct=0)
    nrmse = np.linalg.norm(recon - ref) / np.linalg.norm(ref)
    print("%s weighted+positivity NRMSE vs oracle %.3e, min %.3e" % (kind, nrmse, recon.min()), flush=True)
    assert nrmse <= 1e-3
    # x + alpha * max(-x / (alpha + eps), delta) is >= 0 up to float32 rounding
    assert recon.min() >= -1e-6 * recon.max()
    assert np.allclose(rd['recon_params']['fm_rmse'], stats['fm_rmse'], rtol=1e-3)


def test_prox_map_fixed_point(gpu):
    """mbirjax tests/test_prox.py:31-61: with prox_input and init_recon equal to
    the phantom and a noiseless sinogram, the proximal map stays put."""
    np.random.seed(0)
    m = mj.ParallelBeamModel((32, 40, 128), np.linspace(0.0, np.pi, 32, endpoint=False))
    phantom = m.gen_modified_3d_sl_phantom()
    sino = m.forward_project(phantom)
    m.set_params(verbose=0)
    recon, _ = m.prox_map(phantom, sino, max_iterations=5, init_recon=phantom, print_logs=False, logfile_path=None)
    max_diff = float(np.max(np.abs(np.asarray(phantom) - recon.cpu().numpy())))
    print("prox fixed point max diff %.3e" % max_diff, flush=True)
    assert max_diff < 1e-5


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_prox_map_matches_oracle(gpu, kind):
    m, g, phantom, sino = problem(kind)
    prox_input = (phantom + 0.05 * np.random.default_rng(3).standard_normal(phantom.shape)).astype(np.float32)
    m.set_params(sharpness=0.0, verbose=0, sigma_prox=0.1, sigma_y=0.02, auto_regularize_flag=False,
                 no_warning=True)
    np.random.seed(9)
    recon, _ = m.prox_map(prox_input, sino, max_iterations=3, stop_threshold_change_pct=0, print_logs=False,
                          logfile_path=None)
    om = ovcd.OracleModel(g, sharpness=0.0, sigma_prox=0.1, sigma_y=0.02)
    np.random.seed(9)
    ref, _ = om.prox_map(prox_input, sino, max_iterations=3, stop_threshold_change_pct=0)
    r = recon.cpu().numpy().astype(np.float64)
    nrmse = np.linalg.norm(r - ref) / np.linalg.norm(ref)
    print("%s prox map NRMSE vs oracle %.3e" % (kind, nrmse), flush=True)
    assert nrmse <= 1e-3


def c4_subproblem():
    """The 512^3 benchmark's precision sub-problem: 64 views over 2 pi,
    128 x 192 detector, same magnification (sdd 768, sid 384), 128^3 recon,
    seeded ellipsoid phantom + 1 % noise."""
    V, R, C = 64, 128, 192
    ang = np.linspace(0, 2 * np.pi, V, endpoint=False)
    m = mj.ConeBeamModel((V, R, C), ang, source_detector_dist=768.0, source_iso_dist=384.0)
    m.set_params(recon_shape=(128, 128, 128), sharpness=0.0, verbose=0)
    g = op.Geometry("cone", (V, R, C), ang, 768.0, 384.0, recon_shape=(128, 128, 128),
                    delta_voxel=m.get_params('delta_voxel'))
    phantom = mj.utilities.gen_random_ellipsoid_phantom(g.recon_shape, seed=0)
    ror = np.flatnonzero(opart.get_2d_ror_mask(g.recon_shape))
    sino = op.forward(g, phantom.reshape(-1, 128)[ror], ror)
    sino = (sino + 0.01 * np.max(sino) * np.random.default_rng(1).standard_normal(sino.shape)).astype(np.float32)
    return m, g, phantom, sino


def test_c4_subproblem_precisions_vs_oracle(gpu):
    """10 iterations in each precision mode against the FP64 oracle under the
    same partitions and subset order.  'fp32', 'mixed' (float32 error
    sinogram, bfloat16 delta sinogram) and 'fp16' (float16 error sinogram)
    meet NRMSE <= 1e-3.  'bf16' (bfloat16 error sinogram) is experimental:
    late updates smaller than half a bfloat16 ulp of the residual are rounded
    away, which leaves it at ~3.4e-3 of the oracle (its error against the
    phantom equals fp32's); its line guards that level, not the bar."""
    m, g, phantom, sino = c4_subproblem()
    om = ovcd.OracleModel(g, sharpness=0.0)
    np.random.seed(11)
    ref, _ = om.recon(sino, max_iterations=10, stop_threshold_change_pct=0)
    for precision, bar in (("fp32", 1e-3), ("mixed", 1e-3), ("fp16", 1e-3), ("bf16", 5e-3)):
        np.random.seed(11)
        recon, _ = m.recon(sino, max_iterations=10, stop_threshold_change_pct=0, print_logs=False,
                           logfile_path=None, precision=precision)
        r = recon.cpu().numpy().astype(np.float64)
        nrmse = np.linalg.norm(r - ref) / np.linalg.norm(ref)
        print("C4 sub-problem %s NRMSE vs FP64 oracle %.3e (vs phantom %.3f)" % (
            precision, nrmse, np.linalg.norm(r - phantom) / np.linalg.norm(phantom)), flush=True)
        assert nrmse <= bar


def test_fp16_scaling_is_exact(gpu):
    """The float16 mode solves the problem scaled by a power of two: data
    2^16 times larger (peak ~1.3e6, beyond float16's 65504) reconstructs to
    2^16 times the same recon."""
    m, g, phantom, sino = c4_subproblem()
    out = []
    for f in (1.0, 65536.0):
        np.random.seed(11)
        recon, d = m.recon(sino * np.float32(f), max_iterations=3, stop_threshold_change_pct=0, print_logs=False,
                           logfile_path=None, precision='fp16')
        out.append((recon, d['recon_params']['fm_rmse']))
    assert float(np.max(sino)) * 65536.0 > 65504.0
    diff = float(torch.max(torch.abs(out[1][0] - 65536.0 * out[0][0])) / torch.max(torch.abs(65536.0 * out[0][0])))
    print("fp16 scaled recon: max relative difference %.3e" % diff, flush=True)
    assert torch.isfinite(out[1][0]).all() and diff <= 1e-6
    assert np.allclose(out[0][1], out[1][1], rtol=1e-6)


def test_c1_parallel_64_recon_vs_oracle(gpu):
    """C1: ParallelBeamModel, 64 views over [0, pi), 64 x 64 detector, 64^3
    recon, ellipsoid phantom + 1 % noise, sharpness 0, 10 iterations."""
    V, R, C = 64, 64, 64
    ang = np.linspace(0, np.pi, V, endpoint=False)
    m = mj.ParallelBeamModel((V, R, C), ang)
    g = op.Geometry("parallel", (V, R, C), ang)
    assert m.get_params('recon_shape') == g.recon_shape == (64, 64, 64)
    phantom = mj.utilities.gen_random_ellipsoid_phantom(g.recon_shape, seed=0)
    ror = np.flatnonzero(opart.get_2d_ror_mask(g.recon_shape))
    sino = op.forward(g, phantom.reshape(-1, 64)[ror], ror)
    sino = (sino + 0.01 * np.max(sino) * np.random.default_rng(1).standard_normal(sino.shape)).astype(np.float32)
    m.set_params(sharpness=0.0, verbose=0)
    np.random.seed(0)
    recon, _ = m.recon(sino, max_iterations=10, stop_threshold_change_pct=0, print_logs=False, logfile_path=None)
    np.random.seed(0)
    ref, _ = ovcd.OracleModel(g, sharpness=0.0).recon(sino, max_iterations=10, stop_threshold_change_pct=0)
    r = recon.cpu().numpy().astype(np.float64)
    nrmse = np.linalg.norm(r - ref) / np.linalg.norm(ref)
    print("C1 recon NRMSE vs oracle %.3e, vs phantom %.3f" % (nrmse, np.linalg.norm(r - phantom) /
                                                                np.linalg.norm(phantom)), flush=True)
    assert nrmse <= 1e-3


def test_translation_recon_vs_oracle(gpu):
    """TranslationModel: 5 x 4 grid of x-z translations (mbirjax
    utilities.gen_translation_vectors), 24 x 40 detector, sdd 160, sid 80,
    recon 8 x 44 x 30 with row pitch 1.0, ellipsoid phantom + 1 % noise,
    qggmrf_nbr_wts (0.5, 1, 1), zero initial recon (the geometry has no direct
    recon), 10 iterations."""
    V, R, C = 20, 24, 40
    tv = mj.utilities.gen_translation_vectors(5, 4, 4.0, 3.0)
    m = mj.TranslationModel((V, R, C), tv, source_detector_dist=160.0, source_iso_dist=80.0)
    m.set_params(recon_shape=(8, 44, 30), delta_recon_row=1.0)
    g = op.Geometry("translation", (V, R, C), tv, 160.0, 80.0, recon_shape=(8, 44, 30), delta_recon_row=1.0)
    assert m.get_params('qggmrf_nbr_wts') == (0.5, 1, 1)
    phantom = mj.utilities.gen_random_ellipsoid_phantom(g.recon_shape, seed=0)
    ror = np.flatnonzero(opart.get_2d_ror_mask(g.recon_shape))
    sino = op.forward(g, phantom.reshape(-1, 30)[ror], ror)
    sino = (sino + 0.01 * np.max(sino) * np.random.default_rng(1).standard_normal(sino.shape)).astype(np.float32)
    m.set_params(sharpness=0.0, verbose=0)
    np.random.seed(0)
    with pytest.warns(UserWarning):  # direct recon is not available: zero initial recon
        recon, _ = m.recon(sino, max_iterations=10, stop_threshold_change_pct=0, print_logs=False,
                           logfile_path=None)
    np.random.seed(0)
    ref, _ = ovcd.OracleModel(g, sharpness=0.0, qggmrf_nbr_wts=[0.5, 1.0, 1.0]).recon(
        sino, max_iterations=10, stop_threshold_change_pct=0)
    r = recon.cpu().numpy().astype(np.float64)
    nrmse = np.linalg.norm(r - ref) / np.linalg.norm(ref)
    print("translation recon NRMSE vs oracle %.3e, vs phantom %.3f" % (
        nrmse, np.linalg.norm(r - phantom) / np.linalg.norm(phantom)), flush=True)
    assert nrmse <= 1e-3


def _golden_module(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location(name, os.path.join(os.path.dirname(__file__), "golden",
                                                                     name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_c2_cone_20_iterations_vs_oracle_fixture(gpu):
    """C2: cone 128 views over [0, 2 pi), 128 x 192 detector, sdd 768, sid 384
    (recon 192 x 192 x 128), ellipsoid phantom + 1 % noise, 20 VCD iterations
    of the default schedule.  The FP64 oracle's run of the same 20 iterations
    (about 20 s of host time each) is a committed fixture written by
    tests/golden/make_c2_golden.py: three orthogonal recon slices and the
    fm_rmse, alpha and nmae trajectories."""
    gold = _golden_module("make_c2_golden")
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "c2_cone_vcd20.npz"))
    g, phantom, sino = gold.c2_inputs()
    # the same inputs as the fixture's
    assert abs(np.sum(sino, dtype=np.float64) - fx["sino_sum"]) <= 1e-9 * abs(fx["sino_sum"])
    assert abs(np.sum(sino.astype(np.float64) ** 2) - fx["sino_sumsq"]) <= 1e-9 * fx["sino_sumsq"]
    m = mj.ConeBeamModel(g.sinogram_shape, g.angles, source_detector_dist=gold.SDD, source_iso_dist=gold.SID)
    assert m.get_params('recon_shape') == g.recon_shape == (192, 192, 128)
    m.set_params(sharpness=0.0, verbose=0)
    np.random.seed(0)
    recon, rd = m.recon(sino, max_iterations=gold.ITERATIONS, stop_threshold_change_pct=0, print_logs=False,
                        logfile_path=None)
    r = recon.cpu().numpy().astype(np.float64)
    num = sum(np.sum((a - b.astype(np.float64)) ** 2) for a, b in
              zip(gold.slices(r), (fx["slice_z"], fx["slice_row"], fx["slice_col"])))
    den = sum(np.sum(b.astype(np.float64) ** 2) for b in (fx["slice_z"], fx["slice_row"], fx["slice_col"]))
    nrmse = np.sqrt(num / den)
    p = rd['recon_params']
    fm = np.array(p['fm_rmse'])
    al = np.array(p['alpha_values'])
    fm_gap = np.max(np.abs(fm - fx["fm_rmse"]) / fx["fm_rmse"])
    al_gap = np.max(np.abs(al - fx["alpha_values"]) / fx["alpha_values"])
    print("C2 20 iterations: slices NRMSE vs oracle %.3e, fm_rmse gap %.2e, alpha gap %.2e, recon norm gap %.2e" % (
        nrmse, fm_gap, al_gap, abs(np.linalg.norm(r) - fx["recon_norm"]) / fx["recon_norm"]), flush=True)
    assert len(fm) == gold.ITERATIONS
    assert nrmse <= 1e-5  # the workload's relative-L2 bar
    assert fm_gap <= 1e-5 and al_gap <= 1e-4
    assert abs(np.linalg.norm(r) - fx["recon_norm"]) <= 1e-6 * fx["recon_norm"]


def test_recon_on_a_side_torch_stream(gpu):
    """A recon issued on a non-default torch stream equals the default-stream
    one bitwise.  The first iteration's single subset (the whole region of
    reconstruction over 64 views) has more than 2^18 view-pixel records, so the
    subset update forks its forward binning onto the library's second stream of
    the device that runs the caller's stream, and joins it back."""
    V, R, C = 64, 64, 96
    ang = np.linspace(0, 2 * np.pi, V, endpoint=False)
    m = mj.ConeBeamModel((V, R, C), ang, source_detector_dist=4.0 * C, source_iso_dist=2.0 * C)
    m.set_params(sharpness=0.0, verbose=0)
    rs = m.get_params('recon_shape')
    assert int(opart.get_2d_ror_mask(rs).sum()) * V >= 1 << 18
    phantom = mj.utilities.gen_random_ellipsoid_phantom(rs, seed=3)
    sino = m.forward_project(phantom).cpu().numpy()

    def arr(r):
        return r.detach().cpu().numpy() if torch.is_tensor(r) else np.asarray(r)

    np.random.seed(5)
    r1, _ = m.recon(sino, max_iterations=3, stop_threshold_change_pct=0, print_logs=False, logfile_path=None)
    r1 = arr(r1)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        np.random.seed(5)
        r2, _ = m.recon(sino, max_iterations=3, stop_threshold_change_pct=0, print_logs=False, logfile_path=None)
        r2 = arr(r2)
    torch.cuda.synchronize()
    assert np.array_equal(r1, r2)


def test_padded_subset_duplicates(gpu):
    """A subset that lists pixels more than once (the padding of
    mbirjax/vcd_utils.py:212-217): every copy adds its step, as .at[].add does
    (mbirjax/tomography_model.py:1820-1823), within 1e-5 of the FP64 oracle,
    and bitwise the same on every run -- the copies' steps are equal, so the
    order in which their additions land does not change the sum."""
    m, g, phantom, sino = problem("cone")
    m.set_params(sharpness=0.0, verbose=0)
    m.auto_set_regularization_params(sino)
    rows, cols, nz = g.recon_shape
    ror = np.flatnonzero(opart.get_2d_ror_mask(g.recon_shape))
    base = np.random.default_rng(11).choice(ror, 40, replace=False)
    S = np.concatenate([base, base[:15], base[:5]]).astype(np.int32)  # 5 pixels thrice, 10 twice
    runs = []
    for _ in range(2):
        st = m.vcd_setup(sino, init_recon=0, precision='fp32', max_subset_size=S.size)
        st.update_subset(torch.as_tensor(S, device='cuda'))
        runs.append((st.recon().clone(), st.e.clone()))
    assert torch.equal(runs[0][0], runs[1][0]) and torch.equal(runs[0][1], runs[1][1])
    om = ovcd.OracleModel(g, sharpness=0.0)
    om.params["sigma_y"] = float(m.get_params('sigma_y'))
    om.params["sigma_x"] = float(m.get_params('sigma_x'))
    fm_hess = om.hessian_diagonal().reshape(-1, nz)
    flat, _, _, _ = om.subset_update(np.zeros((rows * cols, nz)), sino.astype(np.float64), S, 1.0, True, fm_hess,
                                     1.0 / om.params["sigma_y"] ** 2, om.qparams(), None, 1.0)
    x = runs[0][0].cpu().numpy().astype(np.float64)
    assert np.any(flat[base[:5]] != 0)
    err = np.linalg.norm(x - flat) / np.linalg.norm(flat)
    print("padded subset update vs oracle %.2e" % err, flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("mode", ["plain", "weighted_positive", "prox"])
def test_planned_subset_updates_bitwise_equal(gpu, mode):
    """Sweeping a partition with each subset's projector setup cached
    (mbx_vcd_subset_update_planned, the plan built on the first sweep) gives
    bit for bit the recon, error sinogram and statistics of sweeps that build
    the setup in every update; also with weights and positivity (whose second
    projection of delta reuses the plan's bins) and for the proximal map."""
    m, g, phantom, sino = problem("cone")
    m.set_params(sharpness=0.0, verbose=0, positivity_flag=(mode == "weighted_positive"))
    w = (np.random.default_rng(5).uniform(0.5, 1.5, sino.shape).astype(np.float32)
         if mode == "weighted_positive" else None)
    m.auto_set_regularization_params(sino, weights=w)
    prox = phantom.astype(np.float32) if mode == "prox" else None
    np.random.seed(4)
    parts = mj.vcd_utils.gen_set_of_pixel_partitions(g.recon_shape, [4, 16])
    nmax = max(int(p.shape[1]) for p in parts)
    states = [m.vcd_setup(sino, weights=w, init_recon=None, prox_input=prox, precision='fp32',
                          max_subset_size=nmax) for _ in range(2)]
    dev = states[0].e.device
    parts = [torch.as_tensor(np.ascontiguousarray(p, dtype=np.int32), device=dev) for p in parts]
    rng = np.random.RandomState(6)
    for k in (0, 1, 1, 0, 1):
        order = rng.permutation(parts[k].shape[0])
        a = states[0].iterate(parts[k], order, cache=True)
        b = states[1].iterate(parts[k], order, cache=False)
        assert a == b
        assert torch.equal(states[0].flat_recon, states[1].flat_recon)
        assert torch.equal(states[0].e, states[1].e)
    assert len(states[0].plans) == 4 + 16 and states[0].plan_bytes > 0 and not states[1].plans
