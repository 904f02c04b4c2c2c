"""mbx_back_project_noalloc (the back projection with its scratch in a
file-scope device array) against mbx_back_project, through the C ABI, on
consecutive calls that share the scratch: a pixel set whose geometry table
exceeds the scratch (projected in smaller chunks: equal to float32 rounding of
the different view-group split), then smaller ones that fit (bitwise equal)."""
import numpy as np
import pytest
import torch

import mbirjax_amd as mj
from mbirjax_amd import _lib

pytestmark = pytest.mark.gpu


def noalloc_back(m, sino_vcr, pix, views=None, coeff_power=1, weights=None):
    nz = m.get_params('recon_shape')[2]
    out = torch.full((pix.numel(), nz), float('nan'), dtype=torch.float32, device=sino_vcr.device)
    nv = sino_vcr.shape[0] if views is None else int(views.numel())
    dt = _lib.DTYPE_F32 if sino_vcr.dtype == torch.float32 else _lib.DTYPE_BF16
    _lib.check(_lib.lib().mbx_back_project_noalloc(
        m._geom, _lib.ptr(m._angles()), _lib.ptr(sino_vcr), dt, int(sino_vcr.shape[0]), _lib.ptr(weights),
        _lib.ptr(pix), int(pix.numel()), _lib.ptr(views), nv, int(coeff_power), 1.0, 0, _lib.ptr(out),
        _lib.stream()))
    return out


def test_noalloc_back_projection_matches_back_project(gpu):
    V, R, C = 128, 128, 192
    m = mj.ConeBeamModel((V, R, C), np.linspace(0, 2 * np.pi, V, endpoint=False), source_detector_dist=768.0,
                         source_iso_dist=384.0)
    rows, cols, nz = m.get_params('recon_shape')
    dev = torch.device('cuda')
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    y = torch.randn((V, C, R), generator=g, device=dev)
    ror = torch.as_tensor(np.flatnonzero(mj.vcd_utils.get_2d_ror_mask((rows, cols))).astype(np.int32), device=dev)
    scratch = int(_lib.lib().mbx_noalloc_scratch_bytes())
    assert ror.numel() * V * 68 > scratch  # the full table does not fit: chunked in the scratch
    print("noalloc call 1: %d pixels x %d views (table %.0f MB > scratch %.0f MB)" % (
        ror.numel(), V, ror.numel() * V * 68 / 1e6, scratch / 1e6), flush=True)
    a = noalloc_back(m, y, ror)
    b = m._back(y, ror)
    torch.cuda.synchronize()
    rel = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    print("  rel-L2 vs mbx_back_project %.3e" % rel, flush=True)
    assert torch.isfinite(a).all() and rel <= 1e-6

    # smaller inputs that fit the scratch: the same plans, bitwise the same sums
    sub = ror[torch.randperm(ror.numel(), generator=torch.Generator().manual_seed(1))[:224].to(dev)]
    print("noalloc call 2: %d pixels" % sub.numel(), flush=True)
    assert torch.equal(noalloc_back(m, y, sub), m._back(y, sub))
    views = torch.as_tensor(np.array([5, 0, 77, 127, 64], np.int32), device=dev)
    y16 = y.to(torch.bfloat16)
    print("noalloc call 3: %d pixels, %d views, bfloat16" % (sub[:50].numel(), views.numel()), flush=True)
    assert torch.equal(noalloc_back(m, y16, sub[:50], views), m._back(y16, sub[:50], views))
    w = torch.rand((V, C, R), generator=g, device=dev)
    print("noalloc call 4: Hessian of 17 pixels with weights", flush=True)
    assert torch.equal(noalloc_back(m, w, sub[:17], coeff_power=2), m._back(w, sub[:17], coeff_power=2))
    torch.cuda.synchronize()
