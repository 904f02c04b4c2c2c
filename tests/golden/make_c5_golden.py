"""Writes tests/golden/c5_sigma_x_fd.npz: FP64 central finite differences of an
unrolled-VCD loss with respect to four per-iteration log(sigma_x) values,
computed with the FP64 oracle only (its projectors and its subset update).

Problem: cone beam, 128 views over [0, 2 pi), a 128 x 192 detector,
source-detector 768 and source-iso 384 (magnification 2), recon 128^3; truth =
clip(U[0,1) - 0.2, 0) from default_rng(7), y = A truth + 0.01 N(0, 1) from the
same generator; sigma_y = 0.1; four full-volume VCD updates (every pixel of the
region of reconstruction in one subset) from x = 0, update i using
sigma_x = exp(log_sigma_x[i]) with the default qGGMRF p, q, T and neighbour
weights; loss = 0.5 ||x_4 - truth||^2.  The derivative of the loss with respect
to each log_sigma_x[i] at log(0.3) is taken by central differences with step
1e-4.  Run once on a CPU host (a few minutes); it is not part of the test run.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import partitions as opart  # noqa: E402
from oracle import projectors as op  # noqa: E402
from oracle import vcd as ovcd  # noqa: E402

OUT = os.path.join(HERE, "c5_sigma_x_fd.npz")
V, R, C = 128, 128, 192
SDD, SID = 768.0, 384.0
SHAPE = (128, 128, 128)
SIGMA_Y, SIGMA_X, UPDATES, STEP = 0.1, 0.3, 4, 1e-4


def c5_inputs():
    """Geometry, truth and data of the problem (deterministic)."""
    ang = np.linspace(0, 2 * np.pi, V, endpoint=False)
    g = op.Geometry("cone", (V, R, C), ang, SDD, SID, recon_shape=SHAPE)
    rng = np.random.default_rng(7)
    truth = np.clip(rng.random(SHAPE) - 0.2, 0, None)
    S = np.flatnonzero(opart.get_2d_ror_mask(SHAPE).reshape(-1)).astype(np.int32)
    y = op.forward(g, truth.reshape(-1, SHAPE[2])[S], S) + 0.01 * rng.standard_normal(g.sinogram_shape)
    return g, truth, y, S


def unrolled_loss(g, truth, y, S, log_sigma_x, om, fm_hess):
    b, _, p, q, T = om.qparams()
    flat = np.zeros((SHAPE[0] * SHAPE[1], SHAPE[2]))
    e = y - op.forward(g, flat[S], S)
    for ls in log_sigma_x:
        flat, e, _, _ = om.subset_update(flat, e, S, 1.0, True, fm_hess, 1.0 / SIGMA_Y ** 2,
                                         (b, float(np.exp(ls)), p, q, T), None, 1.0)
    return 0.5 * float(np.sum((flat.reshape(SHAPE) - truth) ** 2))


def main():
    t0 = time.time()
    g, truth, y, S = c5_inputs()
    om = ovcd.OracleModel(g)
    fm_hess = om.hessian_diagonal().reshape(-1, SHAPE[2])
    ls0 = np.full(UPDATES, np.log(SIGMA_X))
    loss0 = unrolled_loss(g, truth, y, S, ls0, om, fm_hess)
    fd = np.zeros(UPDATES)
    for i in range(UPDATES):
        lp, lm = ls0.copy(), ls0.copy()
        lp[i] += STEP
        lm[i] -= STEP
        fd[i] = (unrolled_loss(g, truth, y, S, lp, om, fm_hess) - unrolled_loss(g, truth, y, S, lm, om, fm_hess)) \
            / (2 * STEP)
        print("log sigma_x[%d]: FD %.9e (%.1fs)" % (i, fd[i], time.time() - t0), flush=True)
    np.savez_compressed(OUT, log_sigma_x=ls0, loss=loss0, fd_grad=fd, step=STEP, sigma_y=SIGMA_Y,
                        delta_voxel=g.delta_voxel, y_sum=np.sum(y), y_sumsq=np.sum(y * y))
    print("wrote %s, loss %.9e" % (OUT, loss0))


if __name__ == "__main__":
    main()
