"""Laws of the cone-beam scanner that fix the projector independently of its formulas.

The scanner (mbirjax/cone_beam.py:646-707): a voxel (i, j, k) sits at world
position x~ = dv (j - (cols-1)/2), y~ = dv (i - (rows-1)/2),
z = dv (k - (Nz-1)/2) + recon_slice_offset.  At view angle theta the scanner
frame has unit vectors e_x = (cos, -sin, 0), e_y = (sin, cos, 0), e_z; the
source is at sid e_y, the flat detector is the plane through
(sid - sdd) e_y spanned by e_x (channels) and e_z (rows), and detector cell
(m, n) has its centre at u = (n - (C-1)/2) dc - det_channel_offset,
v = (m - (R-1)/2) dr - det_row_offset (detector_mn_to_uv, cone_beam.py:709-737).

Everything below is computed from that physical description with plain vector
geometry -- a ray from the source through the voxel centre intersected with
the detector plane, the chord of that ray through the voxel cube, and the
Jacobian of the central projection -- never from the projector's own formulas
(magnification, theta_p, cos(alpha), phi, W).  For the impulse response P(m, n)
of one voxel in one view, the projector must satisfy:

* centroid law: the P-weighted mean cell equals the pinhole image of the voxel
  centre to within 1/(8 W) cells, W the footprint width in cells (a box of
  width W convolved with the unit cell aperture, sampled at cell centres, has
  its discrete centroid within 1/(8 W) of the box centre);
* plateau law: where the footprint is at least two cells wide in both
  directions, max P equals the chord of the central ray through the cube,
  dv / max(|d_x~|, |d_y~|, |d_z|), d the unit ray direction;
* mass law: sum P dc dr equals the voxel volume times the Jacobian of the
  central projection, dv^3 (|Q - S| / |P - S|)^2 / cos(gamma), gamma the ray's
  incidence angle on the detector.

mbirjax's separable footprint (horizontal and vertical fans evaluated apart)
meets the plateau and mass laws to second order in the cone angles: the
relative gap is sqrt(1 + a^2 b^2 / (1 + a^2 + b^2)) - 1 <= a^2 b^2 / 2, with
a = u/sdd and b = v/sdd of the voxel's image, which is the tolerance used.
Shift laws: raising det_channel_offset (det_row_offset) by one pitch moves the
image by exactly one cell towards higher n (m); raising recon_slice_offset by
dv equals moving every voxel up one slice.
"""
import math

import numpy as np


def scanner_frame(geom, theta):
    """(source, detector centre, e_u, e_v, detector normal towards the source) in world coordinates."""
    c, s = math.cos(theta), math.sin(theta)
    ex = np.array([c, -s, 0.0])
    ey = np.array([s, c, 0.0])
    ez = np.array([0.0, 0.0, 1.0])
    sid = geom.source_iso_dist
    sdd = geom.source_detector_dist
    return sid * ey, (sid - sdd) * ey, ex, ez, ey


def voxel_world(geom, i, j, k):
    rows, cols, nz = geom.recon_shape
    dv = geom.delta_voxel
    return np.array([dv * (j - (cols - 1) / 2.0), dv * (i - (rows - 1) / 2.0),
                     dv * (k - (nz - 1) / 2.0) + geom.recon_slice_offset])


def pinhole(geom, theta, P):
    """Image of world point P: fractional (m, n), the ray's detector point
    coordinates (u, v), |Q-S|/|P-S| and cos(gamma)."""
    S, D0, eu, ev, nrm = scanner_frame(geom, theta)
    d = P - S
    t = np.dot(D0 - S, nrm) / np.dot(d, nrm)
    Q = S + t * d
    u = np.dot(Q - D0, eu)
    v = np.dot(Q - D0, ev)
    _, R, C = geom.sinogram_shape
    n = (u + geom.det_channel_offset) / geom.delta_det_channel + (C - 1) / 2.0
    m = (v + geom.det_row_offset) / geom.delta_det_row + (R - 1) / 2.0
    qs = Q - S
    cos_gamma = abs(np.dot(qs, nrm)) / np.linalg.norm(qs)
    return m, n, u, v, np.linalg.norm(qs) / np.linalg.norm(d), cos_gamma


def chord(geom, theta, P):
    """Length of the source ray through the centre of the voxel cube at P."""
    S = scanner_frame(geom, theta)[0]
    d = (P - S) / np.linalg.norm(P - S)
    return geom.delta_voxel / np.max(np.abs(d))


def impulse_responses(project, geom, voxels):
    """{(i, j, k): (V, R, C) response} from project(pixel_index, k) -> (V, R, C)."""
    cols = geom.recon_shape[1]
    return {(i, j, k): np.asarray(project(i * cols + j, k), dtype=np.float64) for (i, j, k) in voxels}


def law_errors(geom, responses, rel_eps=1e-9):
    """Largest violation, relative to each law's tolerance, of the centroid,
    plateau and mass laws over the given impulse responses (a value <= 1
    means the law holds).  rel_eps widens every relative tolerance by the
    arithmetic's rounding (1e-9 for FP64; a float32 projector needs ~1e-5).
    Returns a dict law -> max ratio."""
    dv = geom.delta_voxel
    dc, dr = geom.delta_det_channel, geom.delta_det_row
    worst = dict(centroid=0.0, plateau=0.0, mass=0.0, border=0.0)
    V, R, C = geom.sinogram_shape
    mm, nn = np.meshgrid(np.arange(R), np.arange(C), indexing='ij')
    for (i, j, k), resp in responses.items():
        P = voxel_world(geom, i, j, k)
        for vi in range(V):
            theta = float(geom.angles[vi])
            img = resp[vi]
            tot = img.sum()
            # the laws hold for a footprint the detector holds whole: any
            # response on the outermost cells counts as a violation
            edge = max(np.abs(img[0]).max(), np.abs(img[-1]).max(), np.abs(img[:, 0]).max(),
                       np.abs(img[:, -1]).max())
            worst['border'] = max(worst['border'], np.inf if edge > 0 else 0.0)
            m, n, u, v, jac_mag, cos_gamma = pinhole(geom, theta, P)
            # footprint widths in cells, bounded below without the projector:
            # the voxel's shadow is at least dv * jac_mag * cos(gamma) / sqrt(2) wide
            w_lo = min(dv * jac_mag * cos_gamma / math.sqrt(2) / dc, dv * jac_mag * cos_gamma / dr)
            tol_c = 1.0 / (8.0 * w_lo) + rel_eps * max(R, C)
            if tot <= 0:
                worst['centroid'] = max(worst['centroid'], np.inf)
                continue
            cm = (mm * img).sum() / tot
            cn = (nn * img).sum() / tot
            worst['centroid'] = max(worst['centroid'], abs(cm - m) / tol_c, abs(cn - n) / tol_c)
            a = u / geom.source_detector_dist
            b = v / geom.source_detector_dist
            tol_2 = 0.5 * a * a * b * b + rel_eps
            mass_phys = dv ** 3 * jac_mag ** 2 / cos_gamma / (dc * dr)
            worst['mass'] = max(worst['mass'], abs(tot / mass_phys - 1) / tol_2)
            if w_lo >= 2.0:
                worst['plateau'] = max(worst['plateau'], abs(img.max() / chord(geom, theta, P) - 1) / tol_2)
    return worst


def shift_law_errors(project_full, geom_factory, base, x):
    """Relative violations of the three shift laws.  project_full(geom, x) ->
    (V, R, C) for a full (rows, cols, Nz) volume x; geom_factory(**kw) builds
    the geometry with offsets kw (base: the base offsets)."""
    g0 = geom_factory(**base)
    s0 = project_full(g0, x)
    scale = np.max(np.abs(s0))
    out = {}
    kw = dict(base)
    kw['det_channel_offset'] = base.get('det_channel_offset', 0.0) + g0.delta_det_channel
    s1 = project_full(geom_factory(**kw), x)
    out['channel_offset'] = np.max(np.abs(s1[:, :, 1:] - s0[:, :, :-1])) / scale
    kw = dict(base)
    kw['det_row_offset'] = base.get('det_row_offset', 0.0) + g0.delta_det_row
    s1 = project_full(geom_factory(**kw), x)
    out['row_offset'] = np.max(np.abs(s1[:, 1:, :] - s0[:, :-1, :])) / scale
    kw = dict(base)
    kw['recon_slice_offset'] = base.get('recon_slice_offset', 0.0) + g0.delta_voxel
    s1 = project_full(geom_factory(**kw), x)
    xs = np.zeros_like(x)
    xs[:, :, 1:] = x[:, :, :-1]
    out['slice_offset'] = np.max(np.abs(s1 - project_full(g0, xs))) / scale
    return out
