"""GPU parity of the libmbx projectors against the FP64 oracle.

Bar: relative L2 <= 1e-5 against the FP64 oracle on the same float32 inputs
(the reference's own FP32 projections are pinned only through the oracle: JAX
cannot run here).  Adjoint identity in float32 <= 1e-6 relative.
"""
import numpy as np
import pytest
import torch

import mbirjax_amd as mj
from oracle import projectors as op

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def make_cone(V, R, C, sdd, sid, span=2 * np.pi, **params):
    ang = np.linspace(0, span, V, endpoint=False)
    m = mj.ConeBeamModel((V, R, C), ang, source_detector_dist=sdd, source_iso_dist=sid)
    if params:
        m.set_params(**params)
    g = op.Geometry("cone", (V, R, C), ang, sdd, sid, recon_shape=m.get_params('recon_shape'),
                    delta_det_row=m.get_params('delta_det_row'), delta_det_channel=m.get_params('delta_det_channel'),
                    det_row_offset=m.get_params('det_row_offset'),
                    det_channel_offset=m.get_params('det_channel_offset'), delta_voxel=m.get_params('delta_voxel'),
                    recon_slice_offset=m.get_params('recon_slice_offset'))
    return m, g


def make_par(V, R, C, span=np.pi, **params):
    ang = np.linspace(0, span, V, endpoint=False)
    m = mj.ParallelBeamModel((V, R, C), ang)
    if params:
        m.set_params(**params)
    g = op.Geometry("parallel", (V, R, C), ang, recon_shape=m.get_params('recon_shape'),
                    delta_det_channel=m.get_params('delta_det_channel'),
                    det_channel_offset=m.get_params('det_channel_offset'), delta_voxel=m.get_params('delta_voxel'))
    return m, g


def make_trans(V, R, C, sdd, sid, recon_shape, delta_recon_row, seed=0, tv=None, **params):
    if tv is None:
        rng = np.random.default_rng(seed)
        tv = np.stack([rng.uniform(-3, 3, V), rng.uniform(-2, 2, V), rng.uniform(-2, 2, V)], axis=1)
    m = mj.TranslationModel((V, R, C), tv, source_detector_dist=sdd, source_iso_dist=sid)
    m.set_params(recon_shape=recon_shape, delta_recon_row=delta_recon_row, **params)
    g = op.Geometry("translation", (V, R, C), m.get_params('translation_vectors'), sdd, sid,
                    recon_shape=m.get_params('recon_shape'), delta_det_row=m.get_params('delta_det_row'),
                    delta_det_channel=m.get_params('delta_det_channel'),
                    det_row_offset=m.get_params('det_row_offset'),
                    det_channel_offset=m.get_params('det_channel_offset'), delta_voxel=m.get_params('delta_voxel'),
                    delta_recon_row=m.get_params('delta_recon_row'))
    assert g.psf_radii()[0] == m.get_psf_radius()
    return m, g


CASES = {
    "cone_m2": lambda: make_cone(24, 20, 36, 4 * 36, 2 * 36),
    "cone_m1_ref_test_geometry": lambda: make_cone(16, 10, 32, 4 * 32, 4 * 32, span=np.pi),
    "cone_offsets_odd": lambda: make_cone(13, 70, 45, 200.0, 90.0, det_row_offset=1.3, det_channel_offset=-2.2,
                                          recon_slice_offset=0.7),
    "cone_tall": lambda: make_cone(8, 150, 20, 160.0, 80.0),
    "parallel": lambda: make_par(20, 12, 30),
    "parallel_offset_tall": lambda: make_par(9, 130, 17, det_channel_offset=0.6),
    # more than 1024 slices: several slice chunks plus a partial last one in
    # the back projector, several row blocks in the forward projector
    "parallel_multichunk": lambda: make_par(6, 1100, 10),
    "cone_multichunk": lambda: make_cone(6, 1300, 12, 2600.0, 1300.0, recon_shape=(8, 8, 1200)),
    "translation": lambda: make_trans(12, 20, 36, 144.0, 72.0, (6, 30, 24), 1.5),
    "translation_grid_offsets": lambda: make_trans(20, 24, 40, 160.0, 80.0, (8, 44, 30), 1.0,
                                                   tv=mj.utilities.gen_translation_vectors(5, 4, 4.0, 3.0),
                                                   det_row_offset=0.8, det_channel_offset=-1.1),
    # tap radii 2: a five-channel horizontal window (large voxels), a five-slice
    # vertical window in the general mode (small voxels)
    "cone_psf2": lambda: make_cone(12, 16, 24, 96.0, 48.0, delta_voxel=1.0, recon_shape=(8, 8, 6)),
    "cone_bp2": lambda: make_cone(12, 24, 24, 96.0, 48.0, delta_voxel=0.2, recon_shape=(20, 20, 40)),
    "parallel_psf2": lambda: make_par(10, 8, 20, delta_voxel=2.5, recon_shape=(8, 8, 8)),
    "translation_psf2": lambda: make_trans(12, 20, 36, 144.0, 72.0, (6, 30, 24), 1.5, delta_voxel=1.2),
    # more than 256 detector rows and 449..512 slices: the forward stages its
    # batches through registers; cylinders partly off the detector (spans cut
    # at its edges) and offsets
    "cone_regstage": lambda: make_cone(10, 300, 40, 240.0, 120.0, recon_shape=(12, 12, 480), det_row_offset=2.7,
                                       recon_slice_offset=-1.3),
    "translation_regstage": lambda: make_trans(8, 300, 40, 240.0, 120.0, (6, 12, 470), 1.5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_full_projection_parity(gpu, case):
    print("case", case, flush=True)
    m, g = CASES[case]()
    rows, cols, nz = m.get_params('recon_shape')
    rng = np.random.default_rng(11)
    pix = np.arange(rows * cols, dtype=np.int32)
    x = rng.standard_normal((pix.size, nz)).astype(np.float32)
    s = m.sparse_forward_project(x, pix)
    assert rel(s, op.forward(g, x, pix)) <= TOL
    y = rng.standard_normal(g.sinogram_shape).astype(np.float32)
    b = m.sparse_back_project(y, pix)
    assert rel(b, op.back(g, y, pix)) <= TOL
    h = m.sparse_back_project(np.abs(y), pix, coeff_power=2)
    assert rel(h, op.back(g, np.abs(y), pix, coeff_power=2)) <= TOL


@pytest.mark.parametrize("case", ["cone_m2", "parallel", "cone_offsets_odd", "translation"])
def test_subset_and_view_parity(gpu, case):
    print("case", case, flush=True)
    m, g = CASES[case]()
    rows, cols, nz = m.get_params('recon_shape')
    np.random.seed(3)
    part = mj.gen_pixel_partition(m.get_params('recon_shape'), 7)
    rng = np.random.default_rng(12)
    V = g.sinogram_shape[0]
    views = np.array([V - 1, 0, 2, 2, V // 2], dtype=np.int32)
    for sub in part[:3]:
        x = rng.standard_normal((sub.size, nz)).astype(np.float32)
        assert rel(m.sparse_forward_project(x, sub), op.forward(g, x, sub)) <= TOL
        assert rel(m.sparse_forward_project(x, sub, view_indices=views), op.forward(g, x, sub, views=views)) <= TOL
        y = rng.standard_normal(g.sinogram_shape).astype(np.float32)
        assert rel(m.sparse_back_project(y, sub), op.back(g, y, sub)) <= TOL
        assert rel(m.sparse_back_project(y, sub, view_indices=views), op.back(g, y, sub, views=views)) <= TOL


def test_benchmark_c2_full_projections(gpu):
    """Cone 128 views x 128 x 192, M=2, recon 192x192x128: full forward and back."""
    m, g = make_cone(128, 128, 192, 768.0, 384.0)
    assert m.get_params('recon_shape') == (192, 192, 128)
    phantom = mj.utilities.gen_random_ellipsoid_phantom((192, 192, 128), seed=0)
    np.random.seed(0)
    full = mj.gen_full_indices((192, 192, 128))
    vox = phantom.reshape(-1, 128)[full]
    s = m.sparse_forward_project(vox, full)
    ref = op.forward(g, vox, full)
    r = rel(s, ref)
    print("C2 forward rel-L2 %.3e" % r, flush=True)
    assert r <= TOL
    b = m.sparse_back_project(ref.astype(np.float32), full)
    rb = rel(b, op.back(g, ref.astype(np.float32), full))
    print("C2 back rel-L2 %.3e" % rb, flush=True)
    assert rb <= TOL


def test_benchmark_c4_subset_parity(gpu):
    """The 512^3 benchmark geometry at full size (720 views x 512 x 768, sdd
    3072, sid 1536, recon 512^3): forward and back projection of one K=128
    subset (1602 pixels) over all 720 views against the FP64 oracle (about 20 s
    of host time for the oracle)."""
    ang = np.linspace(0, 2 * np.pi, 720, endpoint=False)
    m = mj.ConeBeamModel((720, 512, 768), ang, source_detector_dist=3072.0, source_iso_dist=1536.0)
    m.set_params(recon_shape=(512, 512, 512))
    g = op.Geometry("cone", (720, 512, 768), ang, 3072.0, 1536.0, recon_shape=(512, 512, 512),
                    delta_voxel=m.get_params('delta_voxel'))
    np.random.seed(0)
    pix = mj.gen_pixel_partition((512, 512, 512), 128)[0]
    assert pix.size == 1602
    rng = np.random.default_rng(21)
    x = rng.random((pix.size, 512)).astype(np.float32)
    r = rel(m.sparse_forward_project(x, pix), op.forward(g, x, pix))
    print("C4 subset forward rel-L2 %.3e" % r, flush=True)
    assert r <= TOL
    y = rng.random((720, 512, 768)).astype(np.float32)
    rb = rel(m.sparse_back_project(y, pix), op.back(g, y, pix))
    print("C4 subset back rel-L2 %.3e" % rb, flush=True)
    assert rb <= TOL


def test_back_projection_wave_layouts(gpu):
    """Subsets of 1024-1791 pixels over 64 or more views back-project in 8 view
    groups with every wave slot of an XCD filled: waves of two pixels and of
    one when the pairs would load some SIMDs one wave more than the rest
    (1100, 1152 and 1400 pixels), pairs only when they are balanced or load
    most SIMDs alike (1280, 1700).  Each against the FP64 oracle."""
    m, g = make_cone(64, 64, 96, 400.0, 200.0, recon_shape=(48, 48, 64))
    from mbirjax_amd import vcd_utils
    ror = np.flatnonzero(vcd_utils.get_2d_ror_mask((48, 48, 64))).astype(np.int32)
    assert ror.size >= 1700
    rng = np.random.default_rng(5)
    pix_all = np.sort(rng.permutation(ror)[:1700]).astype(np.int32)
    y = rng.random((64, 64, 96)).astype(np.float32)
    for n in (1100, 1152, 1280, 1400, 1700):
        pix = pix_all[:n]
        r = rel(m.sparse_back_project(y, pix), op.back(g, y, pix))
        print("back projection of %d pixels rel-L2 %.3e" % (n, r), flush=True)
        assert r <= TOL


@pytest.mark.parametrize("n", [2047, 3071, 4096])
def test_forward_binning_in_parts(gpu, n):
    """Pixel sets of 2047-4096 over 12 views are binned by several blocks per
    view (1, 2 and 4 parts of at least 1024 slots); a 20-channel detector that
    many of the 64 x 64 pixels miss in some views, and part bounds that split
    channel buckets.  Forward projection against the FP64 oracle."""
    m, g = make_cone(12, 16, 20, 96.0, 48.0, recon_shape=(64, 64, 8))
    rng = np.random.default_rng(11)
    pix = np.sort(rng.permutation(64 * 64)[:n]).astype(np.int32)
    x = rng.standard_normal((n, 8)).astype(np.float32)
    r = rel(m.sparse_forward_project(x, pix), op.forward(g, x, pix))
    print("forward projection of %d pixels binned in parts: rel-L2 %.3e" % (n, r), flush=True)
    assert r <= TOL


@pytest.mark.parametrize("case", ["cone_m2", "parallel", "translation"])
def test_adjoint_fp32(gpu, case):
    m, g = CASES[case]()
    rows, cols, nz = m.get_params('recon_shape')
    pix = np.arange(rows * cols, dtype=np.int32)
    rng = np.random.default_rng(13)
    x = torch.as_tensor(rng.standard_normal((pix.size, nz)).astype(np.float32), device='cuda')
    y = torch.as_tensor(rng.standard_normal(g.sinogram_shape).astype(np.float32), device='cuda')
    Ax = m.sparse_forward_project(x, pix)
    Aty = m.sparse_back_project(y, pix)
    lhs = torch.sum(Ax.double() * y.double()).item()
    rhs = torch.sum(Aty.double() * x.double()).item()
    denom = torch.linalg.norm(Ax.double()).item() * torch.linalg.norm(y.double()).item()
    assert abs(lhs - rhs) / denom <= 1e-6


def test_hessian_matches_impulse(gpu):
    """compute_hessian_diagonal == (A^T A e_ijk)_ijk (mbirjax tests/test_projectors.py:264-303)."""
    m, g = CASES["cone_m2"]()
    rows, cols, nz = m.get_params('recon_shape')
    H = m.compute_hessian_diagonal().cpu().numpy()
    i, j, k = rows // 3, cols // 2, nz // 2
    x = np.zeros((rows * cols, nz), np.float32)
    x[i * cols + j, k] = 0.01
    pix = np.arange(rows * cols, dtype=np.int32)
    AtAx = m.sparse_back_project(m.sparse_forward_project(x, pix), pix).cpu().numpy()
    assert np.allclose(H[i, j, k], AtAx[i * cols + j, k] / 0.01, rtol=1e-5)
    Href = op.back(g, np.ones(g.sinogram_shape), pix, coeff_power=2).reshape(rows, cols, nz)
    assert rel(H, Href) <= TOL


@pytest.mark.parametrize("case", ["cone_m2", "cone_multichunk"])  # (1300 rows: whole 256-row column groups)
@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_two_byte_back_projection(gpu, case, dtype):
    """Back projection of a bfloat16 or float16 channel-major sinogram (the
    2-byte error sinograms of precision 'bf16' / 'fp16') against the oracle on
    the same rounded values."""
    print("case", case, dtype, flush=True)
    m, g = CASES[case]()
    rows, cols, nz = m.get_params('recon_shape')
    pix = torch.arange(rows * cols, dtype=torch.int32, device='cuda')
    rng = np.random.default_rng(14)
    y = rng.standard_normal(g.sinogram_shape).astype(np.float32)
    tdt, code = {"bf16": (torch.bfloat16, mj._lib.DTYPE_BF16), "fp16": (torch.float16, mj._lib.DTYPE_F16)}[dtype]
    y2 = torch.as_tensor(y, device='cuda').to(tdt)
    y_rounded = y2.float().cpu().numpy()
    ref = op.back(g, y_rounded, pix.cpu().numpy())
    vcr2 = m.to_channel_major(y2.float(), dtype=code)
    assert vcr2.dtype == tdt and torch.equal(m.from_channel_major(vcr2), y2.float())
    assert rel(m._back(vcr2, pix), ref) <= TOL


def test_transpose_round_trip_and_determinism(gpu):
    m, g = CASES["cone_offsets_odd"]()
    rng = np.random.default_rng(15)
    y = torch.as_tensor(rng.standard_normal(g.sinogram_shape).astype(np.float32), device='cuda')
    back = m.from_channel_major(m.to_channel_major(y))
    assert torch.equal(back, y)
    rows, cols, nz = m.get_params('recon_shape')
    np.random.seed(1)
    sub = mj.gen_pixel_partition(m.get_params('recon_shape'), 4)[1]
    x = rng.standard_normal((sub.size, nz)).astype(np.float32)
    a1 = m.sparse_forward_project(x, sub)
    a2 = m.sparse_forward_project(x, sub)
    assert torch.equal(a1, a2)
    b1 = m.sparse_back_project(y, sub)
    b2 = m.sparse_back_project(y, sub)
    assert torch.equal(b1, b2)


def test_forward_accumulate_and_sumsq(gpu):
    m, g = CASES["cone_m2"]()
    rows, cols, nz = m.get_params('recon_shape')
    pix = torch.arange(rows * cols, dtype=torch.int32, device='cuda')
    rng = np.random.default_rng(16)
    x = torch.as_tensor(rng.standard_normal((rows * cols, nz)).astype(np.float32), device='cuda')
    w = torch.as_tensor(rng.random(g.sinogram_shape[:1] + g.sinogram_shape[2:] + g.sinogram_shape[1:2])
                        .astype(np.float32), device='cuda')
    s, ss = m._fwd(x, pix, layout=mj._lib.LAYOUT_VCR, weights=w, want_sumsq=True)
    assert abs(ss.item() - torch.sum(s.double() ** 2 * w.double()).item()) <= 1e-9 * ss.item()
    s2 = s.clone()
    m._fwd(x, pix, layout=mj._lib.LAYOUT_VCR, out=s2, accumulate=True)
    assert torch.allclose(s2, 2 * s, rtol=1e-6, atol=1e-6)


def test_chunked_full_projection_matches_parts():
    """A pixel set larger than one geometry-table chunk (720 views: about
    11k pixels per chunk) is projected in interleaved chunks; the result must
    equal the sum of the projections of two halves that each fit one chunk, and
    the per-pixel back projection must not depend on the split."""
    from mbirjax_amd import _lib, vcd_utils
    V, R, C = 720, 64, 96
    m = mj.ConeBeamModel((V, R, C), np.linspace(0, 2 * np.pi, V, endpoint=False), source_detector_dist=384.0,
                         source_iso_dist=192.0)
    m.set_params(recon_shape=(128, 128, 16))
    pix = np.flatnonzero(vcd_utils.get_2d_ror_mask((128, 128, 16))).astype(np.int32)
    assert pix.size > 11000  # more than one table chunk at 720 views
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    x = torch.rand((pix.size, 16), device=dev, generator=g)
    p = torch.as_tensor(pix, device=dev)
    h = pix.size // 2
    full = m._fwd(x, p, layout=_lib.LAYOUT_VCR)
    parts = m._fwd(x[:h].contiguous(), p[:h].contiguous(), layout=_lib.LAYOUT_VCR)
    parts += m._fwd(x[h:].contiguous(), p[h:].contiguous(), layout=_lib.LAYOUT_VCR)
    d = float((full.double() - parts.double()).norm() / parts.double().norm())
    assert d <= 1e-6, d
    bp = m._back(full, p)
    bp_parts = torch.cat([m._back(full, p[:h].contiguous()), m._back(full, p[h:].contiguous())])
    assert torch.equal(bp, bp_parts)


def test_forward_channel_block_widths_agree():
    """The channel-major forward narrows its channel blocks (32 -> 16 -> 8)
    while a launch has fewer than 4096 waves.  Here all 512 views run with
    32-channel blocks (512 views x 8 blocks x 1 wave) and an 8-view selection
    with 8-channel blocks; each output element keeps its summation order, so
    the selected views must agree bitwise, and both match the oracle."""
    from mbirjax_amd import _lib
    m, g = make_cone(512, 64, 256, 1024.0, 512.0)
    rs = m.get_params('recon_shape')
    np.random.seed(4)
    sub = mj.gen_pixel_partition(rs, 64)[0]
    rng = np.random.default_rng(14)
    x = rng.standard_normal((sub.size, rs[2])).astype(np.float32)
    views = np.array([0, 77, 128, 255, 256, 301, 444, 511], dtype=np.int32)
    xd = torch.as_tensor(x, device='cuda')
    pd = torch.as_tensor(sub, device='cuda')
    full = m._fwd(xd, pd, layout=_lib.LAYOUT_VCR)
    part = m._fwd(xd, pd, views=torch.as_tensor(views, device='cuda'), layout=_lib.LAYOUT_VCR)
    assert torch.equal(full[torch.as_tensor(views, device='cuda').long()], part)
    ref = op.forward(g, x, sub, views=views)  # (views, R, C)
    assert rel(part.permute(0, 2, 1), ref) <= TOL


def test_chunked_back_projection_small_last_chunk():
    """A back projection over more pixels than one geometry-table chunk (128
    views: 61681 pixels per chunk) whose last chunk is small: every chunk is
    planned as a full one, so the view-group partials stay inside the workspace
    sized for it (a small chunk planned on its own splits its views, which a
    full chunk of this size does not).  The result must match the back
    projections of the two parts, each within one chunk."""
    from mbirjax_amd import vcd_utils
    V, R, C = 128, 4, 416
    m = mj.ParallelBeamModel((V, R, C), np.linspace(0, np.pi, V, endpoint=False))
    m.set_params(recon_shape=(288, 288, 4))
    pix = np.flatnonzero(vcd_utils.get_2d_ror_mask((288, 288, 4))).astype(np.int32)
    chunk = 61681
    assert chunk < pix.size < chunk + 8192, pix.size
    print("pixels", pix.size, flush=True)
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    e = torch.randn((V, C, R), device=dev, generator=g)
    p = torch.as_tensor(pix, device=dev)
    full = m._back(e, p)
    parts = torch.cat([m._back(e, p[:chunk].contiguous()), m._back(e, p[chunk:].contiguous())])
    d = float((full.double() - parts.double()).norm() / parts.double().norm())
    assert d <= 1e-6, d
    assert torch.equal(full[:chunk], parts[:chunk])  # same plan for the first chunk


@pytest.mark.parametrize("case", ["cone_m2", "cone_bp2", "cone_psf2", "parallel", "translation"])
def test_forward_plan_apply_matches_project(gpu, case):
    """mbx_forward_plan + mbx_forward_apply give bitwise what mbx_forward_project
    gives, over repeated applications of one plan (both layouts, accumulate,
    bfloat16 output, a view subset)."""
    from mbirjax_amd import _lib
    print("case", case, flush=True)
    m, g = CASES[case]()
    rows, cols, nz = m.get_params('recon_shape')
    rng = np.random.default_rng(21)
    pix = torch.as_tensor(np.sort(rng.choice(rows * cols, size=max(1, rows * cols // 2), replace=False))
                          .astype(np.int32), device='cuda')
    x1 = torch.as_tensor(rng.standard_normal((pix.numel(), nz)).astype(np.float32), device='cuda')
    x2 = torch.as_tensor(rng.standard_normal((pix.numel(), nz)).astype(np.float32), device='cuda')
    plan = m._fwd_plan(pix)
    for x in (x1, x2):
        assert torch.equal(m._fwd_apply(plan, x), m._fwd(x, pix, layout=_lib.LAYOUT_VCR))
        assert torch.equal(m._fwd_apply(plan, x, layout=_lib.LAYOUT_VRC), m._fwd(x, pix, layout=_lib.LAYOUT_VRC))
    base = m._fwd(x1, pix, layout=_lib.LAYOUT_VCR)
    acc = m._fwd_apply(plan, x2, out=base.clone(), accumulate=True)
    assert torch.equal(acc, m._fwd(x2, pix, layout=_lib.LAYOUT_VCR, out=base.clone(), accumulate=True))
    assert torch.equal(m._fwd_apply(plan, x1, out_dtype=_lib.DTYPE_BF16),
                       m._fwd(x1, pix, layout=_lib.LAYOUT_VCR, out_dtype=_lib.DTYPE_BF16))
    V = g.sinogram_shape[0]
    views = torch.as_tensor(np.arange(1, V, 3, dtype=np.int32), device='cuda')
    vplan = m._fwd_plan(pix, views)
    assert torch.equal(m._fwd_apply(vplan, x1, views), m._fwd(x1, pix, views, layout=_lib.LAYOUT_VCR))


@pytest.mark.parametrize("case", ["cone_m2", "cone_regstage", "cone_bp2", "parallel", "translation"])
def test_back_apply_matches_back_project(gpu, case):
    """mbx_back_apply (the plan's geometry table) gives bitwise what
    mbx_back_project gives: float32, float16 and bfloat16 sinograms, weights,
    coeff_power 2, accumulate and scale, a view subset; and against the oracle."""
    from mbirjax_amd import _lib
    print("case", case, flush=True)
    m, g = CASES[case]()
    rows, cols, nz = m.get_params('recon_shape')
    V, R, C = g.sinogram_shape
    rng = np.random.default_rng(23)
    pix_np = np.sort(rng.choice(rows * cols, size=max(1, rows * cols // 2), replace=False)).astype(np.int32)
    pix = torch.as_tensor(pix_np, device='cuda')
    y = torch.as_tensor(rng.standard_normal((V, C, R)).astype(np.float32), device='cuda')
    w = torch.as_tensor(rng.uniform(0.5, 1.0, (V, C, R)).astype(np.float32), device='cuda')
    plan = m._fwd_plan(pix)
    b = m._back_apply(plan, y)
    assert torch.equal(b, m._back(y, pix))
    assert rel(b, op.back(g, m.from_channel_major(y).cpu().numpy(), pix_np)) <= TOL
    assert torch.equal(m._back_apply(plan, y, weights=w), m._back(y, pix, weights=w))
    assert torch.equal(m._back_apply(plan, y.abs(), coeff_power=2), m._back(y.abs(), pix, coeff_power=2))
    for dt in (torch.float16, torch.bfloat16):
        yd = y.to(dt)
        assert torch.equal(m._back_apply(plan, yd), m._back(yd, pix))
    base = m._back(y, pix)
    assert torch.equal(m._back_apply(plan, y, scale=0.5, out=base.clone(), accumulate=True),
                       m._back(y, pix, scale=0.5, out=base.clone(), accumulate=True))
    views = torch.as_tensor(np.arange(1, V, 3, dtype=np.int32), device='cuda')
    vplan = m._fwd_plan(pix, views)
    assert torch.equal(m._back_apply(vplan, y, views), m._back(y, pix, views))
    # a forward apply after back applies of the same plan is unchanged (the plan is only read)
    x = torch.as_tensor(rng.standard_normal((pix.numel(), nz)).astype(np.float32), device='cuda')
    assert torch.equal(m._fwd_apply(plan, x), m._fwd(x, pix, layout=_lib.LAYOUT_VCR))


@pytest.mark.parametrize("case", ["cone_offsets_odd", "cone_m2", "parallel", "parallel_offset_tall"])
def test_direct_filters_match_oracle(gpu, case):
    """FDK / FBP filtering (cosine weight, ramp over 2C-1 taps, scale) against
    the FP64 oracle's exact Toeplitz restatement."""
    from oracle import filters as ofilt
    print("case", case, flush=True)
    m, g = CASES[case]()
    y = np.random.default_rng(31).standard_normal(g.sinogram_shape).astype(np.float32)
    got = m.direct_filter(y)
    ref = ofilt.fdk_filter(g, y) if case.startswith("cone") else ofilt.fbp_filter(g, y)
    assert rel(got, ref) <= TOL


@pytest.mark.parametrize("C", [48, 2100])  # the Toeplitz-GEMM route, then the rFFT route
def test_ramp_row_convolution_routes(gpu, C):
    from oracle import filters as ofilt
    rng = np.random.default_rng(32)
    rows = rng.standard_normal((3, 5, C)).astype(np.float32)
    h = ofilt.ramp_filter(C)
    got = mj.TomographyModel._ramp_conv_rows(torch.as_tensor(rows, device='cuda'),
                                             torch.as_tensor(h.astype(np.float32), device='cuda'))
    ref = np.stack([[np.convolve(r.astype(np.float64), h, mode='valid') for r in v] for v in rows])
    assert rel(got, ref) <= TOL


def test_out_of_range_device_pixel_indices_contribute_nothing(gpu):
    """Pixel indices on the device are not read back for a range check (no
    host sync): an index outside the recon contributes nothing to the forward
    projection and back-projects to zeros, and the valid pixels' results are
    unchanged.  Host arrays are still checked before the upload."""
    m, g = CASES["cone_m2"]()
    rows, cols, nz = m.get_params('recon_shape')
    rng = np.random.default_rng(21)
    valid = rng.choice(rows * cols, 37, replace=False).astype(np.int32)
    bad = np.array([-5, rows * cols, rows * cols + 1000, -(1 << 30)], dtype=np.int32)
    pix = np.concatenate([valid[:20], bad[:2], valid[20:], bad[2:]])
    x = rng.standard_normal((pix.size, nz)).astype(np.float32)
    keep = np.isin(pix, valid)
    pd = torch.as_tensor(pix, device='cuda')
    a_all = m.sparse_forward_project(torch.as_tensor(x, device='cuda'), pd)
    a_ok = m.sparse_forward_project(torch.as_tensor(x[keep], device='cuda'), torch.as_tensor(pix[keep], device='cuda'))
    assert torch.allclose(a_all, a_ok, rtol=1e-6, atol=1e-6 * float(a_ok.abs().max()))
    y = torch.as_tensor(rng.standard_normal(g.sinogram_shape).astype(np.float32), device='cuda')
    b_all = m.sparse_back_project(y, pd)
    b_ok = m.sparse_back_project(y, torch.as_tensor(pix[keep], device='cuda'))
    assert torch.allclose(b_all[torch.as_tensor(keep, device='cuda')], b_ok, rtol=1e-6,
                          atol=1e-6 * float(b_ok.abs().max()))
    assert torch.count_nonzero(b_all[torch.as_tensor(~keep, device='cuda')]) == 0
    with pytest.raises(ValueError):
        m.sparse_forward_project(x, pix)
