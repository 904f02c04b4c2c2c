"""Writes tests/golden/c2_cone_vcd20.npz: 20 VCD iterations of the FP64 oracle
on the cone-beam workload of 128 views over [0, 2 pi), a 128 x 192 detector,
source-detector 768 and source-iso 384 (magnification 2), recon 192 x 192 x 128,
a seeded random-ellipsoid phantom plus 1 % Gaussian noise (the default
multi-granular schedule, sharpness 0, np.random.seed(0) before the recon as in
the reference's recon).

The fixture holds the per-iteration fm_rmse, alpha and nmae trajectories, three
orthogonal slices of the recon (float32) and checksums of the sinogram, so that
a test can rebuild the same inputs and compare a 20-iteration recon with it.
The oracle needs about 20 s of host time per iteration at this size: run this
once on a CPU host (it is not part of the test run).
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

OUT = os.path.join(HERE, "c2_cone_vcd20.npz")
V, R, C = 128, 128, 192
SDD, SID = 768.0, 384.0
ITERATIONS = 20


def c2_inputs():
    """Geometry, phantom and noisy sinogram of the workload (deterministic)."""
    from mbirjax_amd import utilities  # numpy phantom generator
    ang = np.linspace(0, 2 * np.pi, V, endpoint=False)
    g = op.Geometry("cone", (V, R, C), ang, SDD, SID)
    phantom = utilities.gen_random_ellipsoid_phantom(g.recon_shape, seed=0)
    ror = np.flatnonzero(opart.get_2d_ror_mask(g.recon_shape))
    nz = g.recon_shape[2]
    sino = op.forward(g, phantom.reshape(-1, nz)[ror], ror)
    sino = (sino + 0.01 * np.max(sino) * np.random.default_rng(1).standard_normal(sino.shape)).astype(np.float32)
    return g, phantom, sino


def slices(recon):
    rows, cols, nz = recon.shape
    return (recon[:, :, nz // 2], recon[rows // 2, :, :], recon[:, cols // 2, :])


def main():
    t0 = time.time()
    g, phantom, sino = c2_inputs()
    print("inputs %.1fs" % (time.time() - t0), flush=True)
    np.random.seed(0)
    ref, stats = ovcd.OracleModel(g, sharpness=0.0).recon(sino, max_iterations=ITERATIONS,
                                                          stop_threshold_change_pct=0)
    print("oracle recon %.1fs" % (time.time() - t0), flush=True)
    s0, s1, s2 = (s.astype(np.float32) for s in slices(ref))
    np.savez_compressed(
        OUT, fm_rmse=np.array(stats["fm_rmse"]), alpha_values=np.array(stats["alpha_values"]),
        nmae_update=np.array(stats["nmae_update"]), slice_z=s0, slice_row=s1, slice_col=s2,
        recon_norm=np.linalg.norm(ref), recon_sum=np.sum(ref),
        sino_sum=np.sum(sino, dtype=np.float64), sino_sumsq=np.sum(sino.astype(np.float64) ** 2),
        phantom_nrmse=np.linalg.norm(ref - phantom) / np.linalg.norm(phantom))
    print("wrote %s (phantom NRMSE %.4f)" % (OUT, np.linalg.norm(ref - phantom) / np.linalg.norm(phantom)))


if __name__ == "__main__":
    main()
