"""ORACLE -- test infrastructure only (see oracle/__init__.py).

NumPy FP64 restatement of the qGGMRF prior (mbirjax/qggmrf.py):
  get_2_b_tilde                          qggmrf.py:250-279
  qggmrf_grad_and_hessian_per_cylinder   qggmrf.py:151-188
  qggmrf_grad_and_hessian_per_slice      qggmrf.py:191-247
  qggmrf_gradient_and_hessian_at_indices qggmrf.py:118-148
  b_tilde_by_definition                  qggmrf.py:282-294
  qggmrf_loss                            qggmrf.py:31-65
  compute_qggmrf_grad_and_hessian        qggmrf.py:344-385
  get_b_from_nbr_wts                     qggmrf.py:388-395
  prox_gradient_at_indices               qggmrf.py:8-28

Pinned by (tests/test_oracle.py): the reference's golden grad/hess vectors
(tests/test_qggmrf.py:83-134, committed as tests/golden/qggmrf_golden.json),
2*b_tilde == 2*b_tilde_by_definition, FP64 finite differences of qggmrf_loss
against the gradient, and equality with the full-volume restatement.
"""
import numpy as np

EPS32 = float(np.finfo(np.float32).eps)


def get_b_from_nbr_wts(qggmrf_nbr_wts):
    b = np.array([qggmrf_nbr_wts, qggmrf_nbr_wts], dtype=np.float64).T.flatten()
    b = b / np.sum(b)
    return tuple(float(v) for v in b)


def get_2_b_tilde(delta, b_for_delta, qggmrf_params):
    b, sigma_x, p, q, T = qggmrf_params
    scaled = np.abs(delta) / (T * sigma_x) + EPS32
    d_q2 = scaled ** (q - 2.0)
    d_qp = scaled ** (q - p)
    num = d_q2 * ((q / p) + d_qp)
    den = (1 + d_qp) ** 2
    scale = (T ** (p - 2)) / (sigma_x * sigma_x)
    return b_for_delta * (scale * num / den)


def b_tilde_by_definition(delta, sigma_x, p, q, T):
    a_min = T * sigma_x * EPS32
    abs_delta = np.clip(np.abs(delta), a_min, None)
    ds = abs_delta / (T * sigma_x)
    ds_qp = ds ** (q - p)
    ds_p2 = abs_delta ** (p - 2)
    num = ds_p2 / (2 * sigma_x ** p)
    num = num * ds_qp * ((q / p) + ds_qp)
    return num / (1 + ds_qp) ** 2


def gradient_and_hessian_at_indices(flat_recon, recon_shape, pixel_indices, qggmrf_params):
    """(grad, hess) of shape (len(pixel_indices), Nz), FP64."""
    b, sigma_x, p, q, T = qggmrf_params
    rows, cols, nz = recon_shape[:3]
    x = np.asarray(flat_recon, dtype=np.float64).reshape(rows * cols, nz)
    idx = np.asarray(pixel_indices).reshape(-1)
    cyl = x[idx]
    # cylinder: slice+1 / slice-1 with zero differences at both ends
    delta = np.concatenate([np.zeros((cyl.shape[0], 1)), np.diff(cyl, axis=1), np.zeros((cyl.shape[0], 1))], axis=1)
    bt2 = get_2_b_tilde(delta, 1.0, qggmrf_params)
    bt2d = bt2 * delta
    grad = -b[4] * bt2d[:, 1:] + b[5] * bt2d[:, :-1]
    hess = b[4] * bt2[:, 1:] + b[5] * bt2[:, :-1]
    # in-plane: row+1, row-1, col+1, col-1 with clipped borders
    i, j = np.unravel_index(idx, (rows, cols))
    for (di, dj), bv in zip([(1, 0), (-1, 0), (0, 1), (0, -1)], b[:4]):
        nb = np.clip(i + di, 0, rows - 1) * cols + np.clip(j + dj, 0, cols - 1)
        d = cyl - x[nb]
        t = get_2_b_tilde(d, bv, qggmrf_params)
        grad = grad + t * d
        hess = hess + t
    return grad, hess


def qggmrf_loss(full_recon, qggmrf_params):
    b, sigma_x, p, q, T = qggmrf_params
    b_axis = [(b[j] + b[j + 1]) / (2 * sum(b)) for j in (0, 2, 4)]

    def rho(delta):
        a_min = T * sigma_x * EPS32
        ad = np.clip(np.abs(delta), a_min, None)
        ds = ad / (T * sigma_x)
        dqp = ds ** (q - p)
        return (ad ** p) / (p * sigma_x ** p) * dqp / (1 + dqp)

    loss = 0.0
    for axis in range(3):
        loss += np.sum(b_axis[axis] * rho(np.diff(full_recon, axis=axis)))
    return loss


def full_grad_and_hessian(full_recon, qggmrf_params):
    b, sigma_x, p, q, T = qggmrf_params
    b_axis = [(b[j] + b[j + 1]) / (2 * sum(b)) for j in (0, 2, 4)]
    hess = np.zeros_like(full_recon, dtype=np.float64)
    grad = np.zeros_like(full_recon, dtype=np.float64)
    for axis in range(3):
        shape = list(full_recon.shape)
        shape[axis] = 1
        z = np.zeros(shape)
        d = np.concatenate([z, np.diff(full_recon, axis=axis), z], axis=axis)
        bt = b_tilde_by_definition(d, sigma_x, p, q, T)
        n = bt.shape[axis]
        bp = np.take(bt, range(1, n), axis=axis)
        bm = np.take(bt, range(0, n - 1), axis=axis)
        hess += 2 * b_axis[axis] * (bp + bm)
        dp = np.take(d, range(1, n), axis=axis)
        dm = np.take(d, range(0, n - 1), axis=axis)
        grad += 2 * b_axis[axis] * (-bp * dp + bm * dm)
    return grad, hess


def prox_gradient_at_indices(flat_recon, prox_input, pixel_indices, sigma_prox):
    idx = np.asarray(pixel_indices).reshape(-1)
    return (1.0 / sigma_prox ** 2) * (np.asarray(flat_recon, np.float64)[idx] - np.asarray(prox_input, np.float64)[idx])
