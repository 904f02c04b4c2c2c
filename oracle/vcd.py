"""ORACLE -- test infrastructure only (see oracle/__init__.py).

FP64 restatement of the MBIR/VCD reconstruction driver of mbirjax 0.6.7:
  auto_set_regularization_params   mbirjax/tomography_model.py:756-856
  _get_sino_indicator              mbirjax/tomography_model.py:883-897
  _get_estimate_of_recon_std       mbirjax/tomography_model.py:899-932
  forward_project / back_project   mbirjax/tomography_model.py:531-574
  compute_hessian_diagonal         mbirjax/tomography_model.py:679-708
  recon                            mbirjax/tomography_model.py:966-1084
  vcd_recon                        mbirjax/tomography_model.py:1086-1255
  vcd_partition_iterator           mbirjax/tomography_model.py:1257-1308
  vcd_subset_updater               mbirjax/tomography_model.py:1343-1542
  get_forward_lin_quad             mbirjax/tomography_model.py:1546-1631
  get_forward_model_loss           mbirjax/tomography_model.py:1633-1658
  prox_map                         mbirjax/tomography_model.py:1660-1697
  direct_recon (FBP/FDK)           mbirjax/cone_beam.py:848-871, parallel_beam.py:350-373
The regularisation scalars are computed with float32 NumPy exactly as the
reference does on the host; everything else is FP64.  The global NumPy RNG is
consumed in the reference's order (partitions, then one gen_full_indices per
forward_project/back_project, then one permutation per iteration).
Pinned by (tests/test_oracle.py): the invariant e == y - A x after every
subset, alpha in [eps32, 1.5], the prox fixed point (tests/test_prox.py:31-61)
and the phantom tolerances of tests/test_vcd.py:21-22 at the reference's
test geometry; and (tests/test_oracle_vcd.py) with the quadratic prior
(p = q = 2): the limit equals the direct solution of the MAP normal equations
(A^T W A / sigma_y^2 + L) x = A^T W y / sigma_y^2, weighted and unweighted, the
positivity limit equals the non-negative least-squares minimiser, and one
iteration equals a dense-matrix replay (delta, alpha, recon, error sinogram);
six planted mistakes of the update (PLANTED below) each fail one of these.
The over-relaxation factor of the prior's quadratic term is pinned only by
restatement.
"""
import numpy as np

from . import filters, partitions, projectors, qggmrf

EPS32 = float(np.finfo(np.float32).eps)

DEFAULTS = dict(sigma_y=1.0, sigma_x=1.0, sigma_prox=1.0, p=2.0, q=1.2, T=1.0, qggmrf_nbr_wts=[1.0, 1.0, 1.0],
                auto_regularize_flag=True, positivity_flag=False, snr_db=30.0, sharpness=1.0,
                granularity=[1, 2, 4, 8, 16, 32, 64, 128, 256], partition_sequence=[0, 2, 4, 6, 7])


# Planted restatement mistakes of the subset update, for the tests that show
# the VCD pins catch them (tests/test_oracle_vcd.py); OracleModel.planted is
# None everywhere else.
PLANTED = {
    "prior_sign": "prior gradient with the opposite sign",
    "fm_constant_sigma": "forward-model constant 1/sigma_y instead of 1/sigma_y^2",
    "no_overrelaxation": "prior quadratic term without the over-relaxation factor 1/2",
    "error_update_sign": "error sinogram updated by + alpha A delta",
    "positivity_no_alpha": "positivity clip delta >= -x instead of -x/alpha",
    "hessian_unweighted": "forward Hessian diagonal without the weights",
}


class OracleModel:
    def __init__(self, geom, **params):
        self.geom = geom
        self.params = dict(DEFAULTS)
        self.params.update(params)
        self.check_invariant = False
        self.planted = None

    # ---------------------------------------------------------- projections
    def forward_project(self, recon):
        rows, cols, nz = self.geom.recon_shape
        idx = partitions.gen_full_indices(self.geom.recon_shape)
        vox = np.asarray(recon, np.float64).reshape(-1, nz)[idx]
        return projectors.forward(self.geom, vox, idx)

    def back_project(self, sinogram):
        rows, cols, nz = self.geom.recon_shape
        idx = partitions.gen_full_indices(self.geom.recon_shape)
        cyl = projectors.back(self.geom, sinogram, idx)
        out = np.zeros((rows * cols, nz))
        out[idx] = cyl
        return out.reshape(rows, cols, nz)

    def hessian_diagonal(self, weights=None):
        rows, cols, nz = self.geom.recon_shape
        if weights is None:
            weights = np.ones(self.geom.sinogram_shape)
        h = projectors.back(self.geom, weights, np.arange(rows * cols), coeff_power=2)
        return h.reshape(rows, cols, nz)

    def direct_filter(self, sinogram):
        if self.geom.kind == "cone":
            return filters.fdk_filter(self.geom, sinogram)
        if self.geom.kind == "translation":  # no direct recon: zeros (mbirjax/tomography_model.py:934-966)
            return np.zeros_like(np.asarray(sinogram, np.float64))
        return filters.fbp_filter(self.geom, sinogram)

    def direct_recon(self, sinogram):
        return self.back_project(self.direct_filter(sinogram))

    # ---------------------------------------------------------- regularisation
    def auto_set_regularization_params(self, sinogram, weights=None):
        if self.params["auto_regularize_flag"]:
            V = sinogram.shape[0]
            step = V // np.minimum(20, V)
            small = np.array(sinogram[::step], dtype=np.float32)
            small_w = 1 if weights is None else np.array(weights[::step], dtype=np.float32)
            ind = np.int8(small > (0.01 * 5.0) * np.mean(np.fabs(small)))
            g = self.geom
            mag = g.magnification()
            signal_rms = float(np.average(small_w * small ** 2, None, ind) ** 0.5)
            rel_noise_std = 10 ** (-self.params["snr_db"] / 20)
            rel_pitch = g.delta_voxel / (g.delta_det_channel / mag)
            self.params["sigma_y"] = float(np.float32(rel_noise_std * signal_rms * (rel_pitch ** 0.5)))
            typical = np.average(np.abs(small), weights=ind)
            rows, cols = g.recon_shape[:2]
            path_space = (2 * rows * cols) / (rows + cols) * g.delta_voxel
            path_sino = small.shape[-1] * g.delta_det_channel / mag
            recon_std = typical / np.minimum(path_space, path_sino)
            sx = float(np.float32(0.2 * (2 ** self.params["sharpness"]) * recon_std))
            self.params["sigma_x"] = sx
            self.params["sigma_prox"] = sx
        return {k: float(self.params[k]) for k in ("sigma_y", "sigma_x", "sigma_prox")}

    def qparams(self):
        b = qggmrf.get_b_from_nbr_wts(self.params["qggmrf_nbr_wts"])
        return (b, float(self.params["sigma_x"]), float(self.params["p"]), float(self.params["q"]),
                float(self.params["T"]))

    # ---------------------------------------------------------- VCD
    def recon(self, sinogram, weights=None, init_recon=None, max_iterations=15, stop_threshold_change_pct=0.2,
              first_iteration=0):
        sinogram = np.asarray(sinogram, np.float64)
        reg = self.auto_set_regularization_params(sinogram, weights=weights)
        parts = partitions.gen_set_of_pixel_partitions(self.geom.recon_shape, self.params["granularity"])
        seq = partitions.gen_partition_sequence(self.params["partition_sequence"], max_iterations)[first_iteration:]
        recon, stats = self.vcd_recon(sinogram, parts, seq, stop_threshold_change_pct, weights=weights,
                                      init_recon=init_recon)
        stats["regularization_params"] = reg
        return recon, stats

    def prox_map(self, prox_input, sinogram, weights=None, init_recon=None, stop_threshold_change_pct=0.2,
                 max_iterations=3, first_iteration=0):
        parts = partitions.gen_set_of_pixel_partitions(self.geom.recon_shape, self.params["granularity"])
        seq = partitions.gen_partition_sequence(self.params["partition_sequence"], max_iterations)[first_iteration:]
        return self.vcd_recon(np.asarray(sinogram, np.float64), parts, seq, stop_threshold_change_pct,
                              weights=weights, init_recon=init_recon, prox_input=prox_input)

    def vcd_recon(self, sinogram, parts, seq, stop_threshold_change_pct, weights=None, init_recon=None,
                  prox_input=None):
        rows, cols, nz = self.geom.recon_shape
        const_w = weights is None
        w = 1.0 if const_w else np.asarray(weights, np.float64)
        if init_recon is None:
            init_recon = self.direct_recon(sinogram)
        elif isinstance(init_recon, int) and init_recon == 0:
            init_recon = np.zeros((rows, cols, nz))
        init_recon = np.asarray(init_recon, np.float64)
        e = self.forward_project(init_recon)
        we = w * e
        nrm = np.sum(we * e)
        alpha = np.sum(we * sinogram) / nrm if nrm > 0 else 1.0
        e = sinogram - alpha * e
        flat = (alpha * init_recon).reshape(-1, nz)
        fm_hess = self.hessian_diagonal(None if const_w or self.planted == "hessian_unweighted"
                                        else w).reshape(-1, nz)
        if prox_input is not None:
            prox_input = np.asarray(prox_input, np.float64).reshape(-1, nz)
        sigma_y = float(self.params["sigma_y"])
        fmc = 1.0 / sigma_y ** (1 if self.planted == "fm_constant_sigma" else 2)
        qp = self.qparams()
        sp = float(self.params["sigma_prox"])
        self.y = sinogram
        fm_rmse, nmae, alphas = [], [], []
        for it in range(len(seq)):
            part = parts[seq[it]]
            order = np.random.permutation(part.shape[0])
            ell1 = 0.0
            asum = 0.0
            for si in order:
                S = part[si]
                flat, e, l1, a = self.subset_update(flat, e, S, w, const_w, fm_hess, fmc, qp, prox_input, sp)
                ell1 += l1
                asum += a
            avg_w = 1.0 if const_w else np.average(w)
            fm_rmse.append(float(np.sqrt(fmc * np.mean(e * e * (w / avg_w)))))
            nmae.append(float(ell1 / np.sum(np.abs(flat))))
            alphas.append(asum / part.shape[0])
            if nmae[-1] < stop_threshold_change_pct / 100:
                break
        return flat.reshape(rows, cols, nz), dict(fm_rmse=fm_rmse, nmae_update=nmae, alpha_values=alphas,
                                                  error_sinogram=e)

    def subset_update(self, flat, e, S, w, const_w, fm_hess, fmc, qp, prox_input, sigma_prox):
        rows, cols, nz = self.geom.recon_shape
        if prox_input is None:
            pg, ph = qggmrf.gradient_and_hessian_at_indices(flat, self.geom.recon_shape, S, qp)
        else:
            ph = sigma_prox ** 2
            pg = qggmrf.prox_gradient_at_indices(flat, prox_input, S, sigma_prox)
        if self.planted == "prior_sign":
            pg = -pg
        we = e if const_w else w * e
        fg = -fmc * projectors.back(self.geom, we, S)
        fh = fmc * fm_hess[S]
        delta = -((fg + pg) / (fh + ph))
        pl = np.sum(pg * delta)
        pq = (1.0 if self.planted == "no_overrelaxation" else 0.5) * np.sum(ph * delta ** 2)
        ds = projectors.forward(self.geom, delta, S)
        fl = fmc * np.sum(we * ds)
        fq = fmc * np.sum(ds * ds * w)
        alpha = (fl - pl) / (fq + pq + EPS32)
        alpha = float(np.clip(alpha, EPS32, 1.5))
        if self.params["positivity_flag"]:
            pos = 1.0 if self.planted == "positivity_no_alpha" else 1.0 / (alpha + EPS32)
            delta = np.maximum(-pos * flat[S], delta)
            ds = projectors.forward(self.geom, delta, S)
        flat = flat.copy()
        # every listed copy of a pixel adds its update, as .at[].add does
        # (mbirjax/tomography_model.py:1820-1823): a padded subset can list a
        # pixel twice (mbirjax/vcd_utils.py:212-217)
        np.add.at(flat, S, alpha * delta)
        e = e + alpha * ds if self.planted == "error_update_sign" else e - alpha * ds
        if self.check_invariant:
            idx = np.flatnonzero(partitions.get_2d_ror_mask(self.geom.recon_shape))
            resid = self.y - projectors.forward(self.geom, flat[idx], idx)
            assert np.max(np.abs(resid - e)) <= 1e-9 * max(1.0, np.max(np.abs(self.y)))
        return flat, e, float(np.sum(np.abs(alpha * delta))), alpha
