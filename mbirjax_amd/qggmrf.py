"""qGGMRF prior on the device (mbirjax/qggmrf.py).

qggmrf_gradient_and_hessian_at_indices runs the libmbx kernel
(mbx_qggmrf_grad_hess); the loss is a torch reduction used only for the
optional prior-loss statistic.
"""
import numpy as np
import torch

from . import _lib


def get_b_from_nbr_wts(qggmrf_nbr_wts):
    """[row, col, slice] neighbour weights -> normalised 6-vector in the order
    [row+1, row-1, col+1, col-1, slice+1, slice-1] (mbirjax/qggmrf.py:388-395)."""
    w = np.asarray(qggmrf_nbr_wts, dtype=np.float64)
    b = np.repeat(w, 2)
    return tuple(float(v) for v in b / np.sum(b))


def qggmrf_gradient_and_hessian_at_indices(flat_recon, recon_shape, pixel_indices, qggmrf_params):
    """Surrogate gradient and Hessian diagonal at the cylinders pixel_indices
    (mbirjax/qggmrf.py:118-148); returns two (npix, Nz) float32 device tensors."""
    _lib.require_gpu()
    b, sigma_x, p, q, T = qggmrf_params
    rows, cols, nz = (int(v) for v in recon_shape[:3])
    dev = torch.device('cuda', torch.cuda.current_device())
    x = (flat_recon if isinstance(flat_recon, torch.Tensor) else torch.as_tensor(np.asarray(flat_recon)))
    x = x.detach().to(device=dev, dtype=torch.float32).reshape(rows * cols, nz).contiguous()
    if isinstance(pixel_indices, torch.Tensor):
        pix = pixel_indices.detach().to(device=dev, dtype=torch.int32).reshape(-1).contiguous()
    else:
        pix = torch.as_tensor(_lib.as_i32(pixel_indices), device=dev)
    grad = torch.empty((pix.numel(), nz), dtype=torch.float32, device=dev)
    hess = torch.empty_like(grad)
    qs = _lib.qggmrf_struct(b, sigma_x, p, q, T)
    _lib.check(_lib.lib().mbx_qggmrf_grad_hess(_lib.ptr(x), rows, cols, nz, _lib.ptr(pix), int(pix.numel()), qs,
                                               _lib.ptr(grad), _lib.ptr(hess), _lib.stream()))
    return grad, hess


def qggmrf_loss(full_recon, qggmrf_params):
    """Prior loss sum over neighbour differences of b_axis * rho(delta)
    (mbirjax/qggmrf.py:31-65), in float64 on the tensor's device."""
    b, sigma_x, p, q, T = qggmrf_params
    x = full_recon if isinstance(full_recon, torch.Tensor) else torch.as_tensor(np.asarray(full_recon))
    x = x.to(torch.float64)
    b_axis = [(b[j] + b[j + 1]) / (2 * sum(b)) for j in (0, 2, 4)]
    eps = float(np.finfo(np.float32).eps)
    loss = 0.0
    for axis in range(3):
        d = torch.diff(x, dim=axis)
        ad = torch.clamp(torch.abs(d), min=T * sigma_x * eps)
        ds = ad / (T * sigma_x)
        dqp = ds ** (q - p)
        rho = (ad ** p) / (p * sigma_x ** p) * dqp / (1 + dqp)
        loss += float(b_axis[axis] * torch.sum(rho))
    return loss


def prox_gradient_at_indices(recon, prox_input, pixel_indices, sigma_prox):
    """(1/sigma_prox^2) (x - prox_input) at the cylinders (mbirjax/qggmrf.py:8-28)."""
    idx = pixel_indices.long() if isinstance(pixel_indices, torch.Tensor) else torch.as_tensor(
        np.asarray(pixel_indices), device=recon.device).long()
    return (1.0 / (sigma_prox ** 2.0)) * (recon[idx] - prox_input[idx])
