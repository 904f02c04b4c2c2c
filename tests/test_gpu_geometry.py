"""The scanner laws, the reference's FBP/FDK accuracy test and the quadratic
MAP limit, checked on the HIP path itself (not only through the oracle).

* Scanner laws (tests/geometry_laws.py): the GPU forward projector's impulse
  responses satisfy the centroid, plateau and mass laws and the offset shift
  laws of the physical scanner, on the two pin scanners with footprints of at
  most 4 cells (psf radius 2).  The relative tolerances are widened by 2e-5
  (float32 arithmetic).
* FBP/FDK: mbirjax tests/test_fbp_fdk.py at its size and tolerances through
  ParallelBeamModel / ConeBeamModel forward_project + direct_recon.
* VCD: recon with p = q = 2 reaches the direct solution of the MAP normal
  equations (tests/test_oracle_vcd.py), to float32 accuracy.
"""
import numpy as np
import pytest
import torch

import mbirjax_amd as mj

from geometry_laws import impulse_responses, law_errors, shift_law_errors
from test_oracle_geometry import SCANNERS, VOXELS, make_geom, shift_volume

pytestmark = pytest.mark.gpu

GPU_SCANNERS = ["narrow_fan", "coarse_aniso"]


def gpu_model(name, **override):
    s = dict(SCANNERS[name])
    s.update(override)
    V = s["V"]
    angles = s["theta0"] + np.arange(V) * (2 * np.pi / V)
    m = mj.ConeBeamModel((V, s["R"], s["C"]), angles, source_detector_dist=s["sdd"], source_iso_dist=s["sid"])
    m.set_params(no_warning=True, delta_det_channel=s["delta_det_channel"], delta_det_row=s["delta_det_row"],
                 det_channel_offset=s["det_channel_offset"], det_row_offset=s["det_row_offset"],
                 recon_slice_offset=s["recon_slice_offset"], delta_voxel=s["delta_voxel"],
                 recon_shape=s["recon_shape"])
    return m


@pytest.mark.parametrize("name", GPU_SCANNERS)
def test_gpu_projector_obeys_scanner_laws(gpu, name):
    print("scanner laws on the GPU projector:", name, flush=True)
    geom = make_geom(name)
    m = gpu_model(name)
    assert (m.get_psf_radius(), m.bp_psf_radius) == geom.psf_radii() and geom.psf_radii()[0] <= 2
    nz = geom.recon_shape[2]

    def project(p, k):
        x = np.zeros((1, nz), np.float32)
        x[0, k] = 1.0
        return m.sparse_forward_project(x, np.array([p], np.int32)).cpu().numpy()

    errs = law_errors(geom, impulse_responses(project, geom, VOXELS[name]), rel_eps=2e-5)

    def project_full(g, x):
        kw = {k: getattr(g, k) for k in ("det_channel_offset", "det_row_offset", "recon_slice_offset")}
        mm = gpu_model(name, **kw)
        rows, cols, _ = g.recon_shape
        return mm.sparse_forward_project(x.reshape(-1, nz).astype(np.float32),
                                         np.arange(rows * cols, dtype=np.int32)).cpu().numpy().astype(np.float64)

    base = {k: SCANNERS[name][k] for k in ("det_channel_offset", "det_row_offset", "recon_slice_offset")}
    shift = shift_law_errors(project_full, lambda **kw: make_geom(name, **kw), base, shift_volume(geom))
    errs.update({k: v / 2e-6 for k, v in shift.items()})
    print(name, {k: "%.3g" % v for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= 1.0}
    assert not bad, bad


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_gpu_fbp_fdk_reference_tolerances(gpu, kind):
    """mbirjax tests/test_fbp_fdk.py:14-79 on the HIP path."""
    print("FBP/FDK reference tolerances:", kind, flush=True)
    V = R = C = 128
    if kind == "cone":
        model = mj.ConeBeamModel((V, R, C), np.linspace(-np.pi, np.pi, V, endpoint=False),
                                 source_detector_dist=4 * C, source_iso_dist=2 * C)
        tol = {'nrmse': 0.27, 'max_diff': 0.51, 'pct_95': 0.1}
    else:
        model = mj.ParallelBeamModel((V, R, C), np.linspace(-np.pi / 2, np.pi / 2, V, endpoint=False))
        tol = {'nrmse': 0.21, 'max_diff': 0.41, 'pct_95': 0.064}
    phantom = model.gen_modified_3d_sl_phantom()
    sino = model.forward_project(phantom)
    recon = model.direct_recon(sino, filter_name="ramp").cpu().numpy().astype(np.float64)
    ph = np.asarray(phantom, np.float64)
    diff = np.abs(recon - ph)
    stats = {'nrmse': np.linalg.norm(recon - ph) / np.linalg.norm(ph), 'max_diff': diff.max(),
             'pct_95': np.percentile(diff, 95)}
    print(kind, {k: "%.4f" % v for k, v in stats.items()}, flush=True)
    assert all(stats[k] < tol[k] for k in tol), stats


@pytest.mark.parametrize("positivity", [False, True])
def test_gpu_vcd_reaches_quadratic_map(gpu, positivity):
    from test_oracle_vcd import NBR_WTS, SIGMA_X, SIGMA_Y, Problem
    print("VCD quadratic MAP limit on the GPU, positivity", positivity, flush=True)
    prob = Problem(sparse_truth=positivity)
    g = prob.geom
    m = mj.ConeBeamModel(g.sinogram_shape, g.angles, source_detector_dist=g.source_detector_dist,
                         source_iso_dist=g.source_iso_dist)
    assert m.get_params('recon_shape') == g.recon_shape
    m.set_params(no_warning=True, p=2.0, q=2.0, sigma_y=SIGMA_Y, sigma_x=SIGMA_X, auto_regularize_flag=False,
                 granularity=[1, 4, 16], partition_sequence=[0, 1, 2], qggmrf_nbr_wts=list(NBR_WTS),
                 positivity_flag=positivity, verbose=0)
    np.random.seed(0)
    recon, _ = m.recon(prob.y.reshape(g.sinogram_shape).astype(np.float32), init_recon=0, max_iterations=60,
                       stop_threshold_change_pct=0, print_logs=False, logfile_path=None)
    x = recon.cpu().numpy().astype(np.float64).reshape(-1)
    xs = prob.map_estimate(positivity)
    err = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    print("relative distance to the MAP %.3e" % err, flush=True)
    assert err < 5e-5
