"""Differentiable projectors and unrolled VCD on the libmbx kernels.

The reference's VCD is not differentiable (it syncs alpha to the host and
draws the subset order from NumPy, mbirjax/tomography_model.py:1281, 1528).
This module adds:

* ``ForwardProjector`` / ``BackProjector``: ``torch.autograd.Function`` s whose
  forward is the hand-written forward (back) projector and whose backward is the
  hand-written back (forward) projector, i.e. the exact adjoint pair of
  ``mbx_forward_project`` / ``mbx_back_project``.
* ``forward_project`` / ``back_project``: differentiable versions of
  ``TomographyModel.forward_project`` / ``back_project`` (same region of
  reconstruction, same outputs).
* ``qggmrf_grad_hess_full``: the 6-neighbour qGGMRF surrogate gradient and
  Hessian (mbirjax/qggmrf.py:151-279) as torch operations, differentiable in the
  recon and in sigma_x.
* ``UnrolledVCD``: a fixed number of full-volume VCD updates
  (mbirjax/tomography_model.py:1343-1542 with a single subset) with a learnable
  sigma_x per iteration; the error sinogram and recon are checkpointed every
  ``checkpoint_every`` iterations and everything in between is recomputed in the
  backward pass (``torch.utils.checkpoint``).
"""
import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import _lib
from . import vcd_utils

EPS32 = float(np.finfo(np.float32).eps)


def _ror_indices(model, device):
    """All region-of-reconstruction pixels, ascending (the pixels
    TomographyModel.forward_project uses), without drawing from the RNG."""
    recon_shape = model.get_params('recon_shape')
    if model.use_ror_mask:
        idx = np.flatnonzero(vcd_utils.get_2d_ror_mask(recon_shape).reshape(-1))
    else:
        idx = np.arange(recon_shape[0] * recon_shape[1])
    return torch.as_tensor(idx.astype(np.int32), device=device)


def _layout_code(layout):
    if layout == 'vrc':
        return _lib.LAYOUT_VRC
    if layout == 'vcr':
        return _lib.LAYOUT_VCR
    raise ValueError("layout must be 'vrc' (views, rows, channels) or 'vcr' (views, channels, rows)")


def _project(model, vox, pix, layout):
    code = _layout_code(layout)
    return model._fwd_rows(vox, pix) if code == _lib.LAYOUT_VRC else model._fwd(vox, pix, layout=code)


class ForwardProjector(torch.autograd.Function):
    """y = A x for cylinders x (npix, Nz) at pixels pix; backward g -> A^T g."""

    @staticmethod
    def forward(ctx, vox, model, pix, layout):
        ctx.model, ctx.layout = model, layout
        ctx.save_for_backward(pix)
        return _project(model, vox.detach().to(torch.float32).contiguous(), pix, layout)

    @staticmethod
    def backward(ctx, g):
        (pix,) = ctx.saved_tensors
        model = ctx.model
        g = g.detach().to(torch.float32).contiguous()
        if ctx.layout == 'vrc':
            g = model._transpose(g, _lib.LAYOUT_VRC, _lib.LAYOUT_VCR)
        return model._back(g, pix), None, None, None


class BackProjector(torch.autograd.Function):
    """x = A^T y for a sinogram y at pixels pix; backward g -> A g."""

    @staticmethod
    def forward(ctx, sino, model, pix, layout):
        ctx.model, ctx.layout = model, layout
        ctx.save_for_backward(pix)
        s = sino.detach().to(torch.float32).contiguous()
        if layout == 'vrc':
            s = model._transpose(s, _lib.LAYOUT_VRC, _lib.LAYOUT_VCR)
        return model._back(s, pix)

    @staticmethod
    def backward(ctx, g):
        (pix,) = ctx.saved_tensors
        out = _project(ctx.model, g.detach().to(torch.float32).contiguous(), pix, ctx.layout)
        return out, None, None, None


def forward_project(model, recon, pixel_indices=None, layout='vrc'):
    """Differentiable forward projection.

    recon: (rows, cols, Nz) tensor, projected over the region of reconstruction
    like TomographyModel.forward_project; or, with pixel_indices, the cylinders
    (npix, Nz) at those pixels like sparse_forward_project.  Returns
    (views, rows, channels) (layout='vrc') or (views, channels, rows) ('vcr').
    """
    _lib.require_gpu()
    if pixel_indices is None:
        rows, cols, nz = model.get_params('recon_shape')
        if tuple(recon.shape) != (rows, cols, nz):
            raise ValueError('recon must have shape {}'.format((rows, cols, nz)))
        pix = _ror_indices(model, recon.device)
        vox = recon.reshape(rows * cols, nz)[pix.long()]
    else:
        pix = model._indices(pixel_indices, device_check=False)
        vox = recon
    return ForwardProjector.apply(vox, model, pix, layout)


def back_project(model, sinogram, pixel_indices=None, layout='vrc'):
    """Differentiable back projection (adjoint of forward_project): a
    (rows, cols, Nz) recon, or the cylinders (npix, Nz) at pixel_indices."""
    _lib.require_gpu()
    if pixel_indices is None:
        rows, cols, nz = model.get_params('recon_shape')
        pix = _ror_indices(model, sinogram.device)
        cyl = BackProjector.apply(sinogram, model, pix, layout)
        full = torch.zeros((rows * cols, nz), dtype=cyl.dtype, device=cyl.device)
        full = full.index_copy(0, pix.long(), cyl)
        return full.reshape(rows, cols, nz)
    return BackProjector.apply(sinogram, model, model._indices(pixel_indices, device_check=False), layout)


# ---------------------------------------------------------------------- qGGMRF in torch
def _two_b_tilde(delta, b_r, sigma_x, p, q, T):
    """2 b~(delta) (mbirjax/qggmrf.py:250-279)."""
    s = torch.abs(delta) / (T * sigma_x) + EPS32
    d_q2 = s ** (q - 2.0)
    d_qp = s ** (q - p)
    return b_r * (T ** (p - 2.0) / (sigma_x * sigma_x)) * d_q2 * ((q / p) + d_qp) / (1.0 + d_qp) ** 2


def qggmrf_grad_hess_full(x, b, sigma_x, p, q, T):
    """Surrogate gradient and Hessian of the qGGMRF prior at every voxel of
    x (rows, cols, Nz): neighbours row+-1, col+-1 clipped to the voxel itself at
    the borders, slice+-1 with zero difference at the ends
    (mbirjax/qggmrf.py:151-247).  Differentiable in x and sigma_x."""
    grad = torch.zeros_like(x)
    hess = torch.zeros_like(x)
    for axis, (bp, bm) in enumerate(((b[0], b[1]), (b[2], b[3]), (b[4], b[5]))):
        n = x.shape[axis]
        idx_p = torch.clamp(torch.arange(n, device=x.device) + 1, max=n - 1)
        idx_m = torch.clamp(torch.arange(n, device=x.device) - 1, min=0)
        for bv, idx in ((bp, idx_p), (bm, idx_m)):
            d = x - torch.index_select(x, axis, idx)
            t = _two_b_tilde(d, bv, sigma_x, p, q, T)
            grad = grad + t * d
            hess = hess + t
    return grad, hess


# ---------------------------------------------------------------------- unrolled VCD
def vcd_full_update(x, e, pix, w, fm_hess, fmc, qparams, project, backproject, positivity):
    """One VCD update over the single subset `pix` (every ROR pixel)
    (mbirjax/tomography_model.py:1343-1542), written with differentiable
    operations.  x: (rows, cols, Nz); e: error sinogram; w: weights or None;
    fm_hess: (npix, Nz) diag(A^T W A) at pix; project/backproject: cylinders
    <-> sinogram callables.  Cost terms and alpha are formed in float64."""
    b, sigma_x, p, q, T = qparams
    rows, cols, nz = x.shape
    flat = x.reshape(rows * cols, nz)
    li = pix.long()
    pg_full, ph_full = qggmrf_grad_hess_full(x, b, sigma_x, p, q, T)
    pg = pg_full.reshape(rows * cols, nz)[li]
    ph = ph_full.reshape(rows * cols, nz)[li]
    we = e if w is None else w * e
    fg = -fmc * backproject(we)
    fh = fmc * fm_hess
    delta = -((fg + pg) / (fh + ph))
    pl = torch.sum(pg.double() * delta.double())
    pq = 0.5 * torch.sum(ph.double() * delta.double() ** 2)
    ad = project(delta)
    fl = fmc * torch.sum(we.double() * ad.double())
    fq = fmc * torch.sum((ad.double() ** 2) if w is None else (w.double() * ad.double() ** 2))
    alpha = torch.clamp((fl - pl) / (fq + pq + EPS32), EPS32, 1.5)
    a = alpha.to(e.dtype)  # float32 on the device, like e -= fl32(alpha) * A delta in the kernels
    if positivity:
        delta = torch.maximum(-(1.0 / (a + EPS32)) * flat[li], delta)
        ad = project(delta)
    flat = flat.index_add(0, li, a * delta)
    e = e - a * ad
    return flat.reshape(rows, cols, nz), e


class UnrolledVCD(torch.nn.Module):
    """num_iterations full-volume VCD updates of a cone- or parallel-beam model
    with a learnable sigma_x per iteration (initialised from the model's
    sigma_x).  The error sinogram is kept channel-major (views, channels, rows)
    on the device.  With checkpoint_every=k, only (recon, error sinogram) at
    every k-th iteration are stored for the backward pass and the iterations in
    between are recomputed."""

    def __init__(self, model, num_iterations, checkpoint_every=2):
        super().__init__()
        _lib.require_gpu()
        self.model = model
        self.num_iterations = int(num_iterations)
        self.checkpoint_every = checkpoint_every
        sx = float(model.get_params('sigma_x'))
        self.log_sigma_x = torch.nn.Parameter(torch.full((self.num_iterations,), float(np.log(sx)),
                                                         dtype=torch.float64))

    def _qparams(self, sigma_x):
        m = self.model
        b = m._qggmrf_params()[0]
        return b, sigma_x, float(m.get_params('p')), float(m.get_params('q')), float(m.get_params('T'))

    def forward(self, sinogram, init_recon, weights=None):
        """sinogram: (views, rows, channels); init_recon: (rows, cols, Nz).
        Returns the recon after num_iterations updates."""
        m = self.model
        dev = init_recon.device
        pix = _ror_indices(m, dev)
        y = m.to_channel_major(sinogram)
        w = None if weights is None else m.to_channel_major(weights)
        fm_hess = m._hessian_vcr(w if w is not None else torch.ones_like(y))[pix.long()]  # (npix, Nz)
        fmc = 1.0 / float(m.get_params('sigma_y')) ** 2
        positivity = bool(m.get_params('positivity_flag'))

        def project(vox):
            return ForwardProjector.apply(vox, m, pix, 'vcr')

        def backproject(s):
            return BackProjector.apply(s, m, pix, 'vcr')

        rows, cols, nz = init_recon.shape
        x = init_recon
        e = y - project(x.reshape(rows * cols, nz)[pix.long()])

        def run(it0, it1, x, e):
            for it in range(it0, it1):
                sigma_x = torch.exp(self.log_sigma_x[it]).to(torch.float32)
                x, e = vcd_full_update(x, e, pix, w, fm_hess, fmc, self._qparams(sigma_x), project, backproject,
                                       positivity)
            return x, e

        k = self.checkpoint_every
        if not k or not torch.is_grad_enabled():
            x, e = run(0, self.num_iterations, x, e)
            return x
        for it0 in range(0, self.num_iterations, k):
            it1 = min(self.num_iterations, it0 + k)
            x, e = checkpoint(run, it0, it1, x, e, use_reentrant=False)
        return x
