"""Pins the cone-beam oracle against the scanner's geometry (no GPU).

The laws (tests/geometry_laws.py) come from the physical scanner -- pinhole
image, chord through the voxel, Jacobian of the central projection, offset
shifts -- not from the projector's formulas, so a mistake made alike in the
oracle and the kernels (which the adjoint and dense-matrix checks cannot see)
fails here.  The last test plants each such mistake in the oracle and shows
that at least one law catches it.
"""
import numpy as np
import pytest

from oracle import projectors as op

from geometry_laws import impulse_responses, law_errors, shift_law_errors

# Three scanners: fine detector pitch (footprints >= 2 cells wide, so the
# plateau law applies) with offsets on every axis and cone angles to ~0.35 rad;
# a narrow fan with the same plateau property and offsets but footprints of at
# most 4 cells (psf radius 2, which the GPU projectors also take); and a
# coarser, anisotropic one without offsets.  Angles avoid the axes.
SCANNERS = {
    "narrow_fan": dict(V=8, R=48, C=48, sdd=120.0, sid=60.0, recon_shape=(6, 6, 10), delta_voxel=1.6,
                       delta_det_channel=1.0, delta_det_row=1.0, det_channel_offset=-1.7, det_row_offset=2.2,
                       recon_slice_offset=-0.9, theta0=0.7),
    "fine_offsets": dict(V=8, R=64, C=60, sdd=60.0, sid=30.0, recon_shape=(6, 6, 8), delta_voxel=2.0,
                         delta_det_channel=1.0, delta_det_row=1.0, det_channel_offset=2.7, det_row_offset=-1.9,
                         recon_slice_offset=1.1, theta0=0.3),
    "coarse_aniso": dict(V=6, R=40, C=36, sdd=80.0, sid=32.0, recon_shape=(7, 5, 9), delta_voxel=1.0,
                         delta_det_channel=1.3, delta_det_row=0.9, det_channel_offset=0.0, det_row_offset=0.0,
                         recon_slice_offset=0.0, theta0=-0.45),
}

VOXELS = {
    "narrow_fan": [(0, 0, 0), (5, 5, 9), (2, 3, 4), (5, 0, 2), (0, 4, 7), (3, 2, 5)],
    "fine_offsets": [(0, 0, 0), (5, 5, 7), (2, 3, 4), (5, 0, 1), (0, 4, 6), (3, 2, 3)],
    "coarse_aniso": [(0, 0, 0), (6, 4, 8), (3, 2, 4), (6, 0, 2), (1, 4, 7)],
}


def make_geom(name, **override):
    s = dict(SCANNERS[name])
    s.update(override)
    V = s.pop("V")
    theta0 = s.pop("theta0")
    angles = theta0 + np.arange(V) * (2 * np.pi / V)
    return op.Geometry("cone", (V, s.pop("R"), s.pop("C")), angles, s.pop("sdd"), s.pop("sid"), **s)


def oracle_impulses(geom, voxels):
    nz = geom.recon_shape[2]

    def project(p, k):
        x = np.zeros((1, nz))
        x[0, k] = 1.0
        return op.forward(geom, x, np.array([p]))
    return impulse_responses(project, geom, voxels)


def oracle_full(geom, x):
    rows, cols, nz = geom.recon_shape
    return op.forward(geom, x.reshape(-1, nz), np.arange(rows * cols))


def shift_volume(geom, seed=3):
    x = np.random.default_rng(seed).random(geom.recon_shape)
    x[:, :, -1] = 0.0  # the slice-offset law moves every voxel up one slice
    return x


def all_law_errors(name):
    geom = make_geom(name)
    errs = law_errors(geom, oracle_impulses(geom, VOXELS[name]))
    base = {k: SCANNERS[name][k] for k in ("det_channel_offset", "det_row_offset", "recon_slice_offset")}
    shift = shift_law_errors(oracle_full, lambda **kw: make_geom(name, **kw), base, shift_volume(geom))
    errs.update({k: v / 1e-9 for k, v in shift.items()})  # shift laws hold to rounding
    return errs


@pytest.mark.parametrize("name", sorted(SCANNERS))
def test_scanner_is_valid_for_the_laws(name):
    """The pin scanners keep every footprint inside the detector and, for the
    fine one, at least two cells wide (so the plateau law is exercised)."""
    geom = make_geom(name)
    errs = law_errors(geom, oracle_impulses(geom, VOXELS[name]))
    assert errs["border"] == 0.0
    if name in ("fine_offsets", "narrow_fan"):
        assert errs["plateau"] > 0.0
    if name == "narrow_fan":
        assert geom.psf_radii() == (2, 1)


@pytest.mark.parametrize("name", sorted(SCANNERS))
def test_oracle_obeys_scanner_laws(name):
    errs = all_law_errors(name)
    bad = {k: v for k, v in errs.items() if not v <= 1.0}
    assert not bad, bad


MISTAKES = {
    1: "1/cos(phi) dropped",
    2: "detector offset signs flipped",
    3: "W_r without the magnification",
    4: "recon_slice_offset sign flipped",
    5: "cos(alpha) from theta + theta_p",
    6: "1/cos(theta_p) in W_c dropped",
    7: "magnification 1/(1/M + y/sdd)",
    8: "view rotation by -theta",
}


@pytest.mark.parametrize("number", sorted(MISTAKES))
def test_planted_mistake_breaks_a_law(number):
    """Each planted restatement mistake violates at least one law on at least
    one scanner, while forward and back stay an exact adjoint pair."""
    violated = []
    with op.planted_mistake(number):
        geom = make_geom("fine_offsets")
        rng = np.random.default_rng(5)
        rows, cols, nz = geom.recon_shape
        pix = np.arange(rows * cols)
        x = rng.standard_normal((pix.size, nz))
        y = rng.standard_normal(geom.sinogram_shape)
        lhs = np.sum(y * op.forward(geom, x, pix))
        assert abs(lhs - np.sum(op.back(geom, y, pix) * x)) <= 1e-11 * (abs(lhs) + 1)
        for name in sorted(SCANNERS):
            errs = all_law_errors(name)
            violated += ["%s:%s=%.3g" % (name, k, v) for k, v in errs.items() if not v <= 1.0]
    assert violated, MISTAKES[number]


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_oracle_fbp_fdk_reference_tolerances(kind):
    """The reference's own direct-recon accuracy test (mbirjax
    tests/test_fbp_fdk.py:14-79): a 128^3 low-dynamic-range Shepp-Logan
    phantom, 128 views x 128 x 128, forward projected and reconstructed by
    FBP (parallel, views over [-pi/2, pi/2)) or FDK (cone, magnification 2,
    views over [-pi, pi)), within the reference's tolerances.  The
    reconstruction inverts the physical ray transform, so it fixes the
    projector's overall scale (a footprint width without the magnification
    fails it)."""
    from mbirjax_amd import utilities
    from oracle import vcd as ovcd
    V = R = C = 128
    if kind == "cone":
        geom = op.Geometry("cone", (V, R, C), np.linspace(-np.pi, np.pi, V, endpoint=False), 4 * C, 2 * C)
        tol = {'nrmse': 0.27, 'max_diff': 0.51, 'pct_95': 0.1}
    else:
        geom = op.Geometry("parallel", (V, R, C), np.linspace(-np.pi / 2, np.pi / 2, V, endpoint=False))
        tol = {'nrmse': 0.21, 'max_diff': 0.41, 'pct_95': 0.064}
    phantom = utilities.generate_3d_shepp_logan_low_dynamic_range(geom.recon_shape).astype(np.float64)
    om = ovcd.OracleModel(geom)
    np.random.seed(0)
    recon = om.direct_recon(om.forward_project(phantom))
    diff = np.abs(recon - phantom)
    stats = {'nrmse': np.linalg.norm(recon - phantom) / np.linalg.norm(phantom), 'max_diff': diff.max(),
             'pct_95': np.percentile(diff, 95)}
    assert all(stats[k] < tol[k] for k in tol), stats
