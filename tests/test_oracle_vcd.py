"""Pins the VCD oracle with the quadratic prior (no GPU).

With p = q = 2 the qGGMRF surrogate weight is the constant 2 b~ = b / (2 sigma_x^2),
so the prior is the quadratic 1/2 x^T L x, L the graph Laplacian of the
6-neighbourhood with edge weight b_axis / (2 sigma_x^2) (clipped in-plane
borders and reflected slice ends add no edge), and the MAP cost is

    f(x) = (1 / (2 sigma_y^2)) (y - A x)^T W (y - A x) + 1/2 x^T L x

over the region-of-reconstruction pixels (the others stay 0).  Its minimiser
solves (A^T W A / sigma_y^2 + L) x = A^T W y / sigma_y^2 restricted to those
pixels; with positivity it is the non-negative least-squares solution of the
same quadratic.  A is the brute-force dense matrix of the projector (pinned by
tests/test_oracle.py and tests/test_oracle_geometry.py) and L is built here
from the neighbourhood alone, so the fixed-point checks pin the VCD update's
gradient, its forward-model constant and the error-sinogram bookkeeping
(mbirjax/tomography_model.py:1343-1542) by linear algebra.

The step size alpha and the curvature in the denominator of delta only set the
speed, not the fixed point.  They are checked against a dense replay of one
iteration: delta = -grad / (diag(A^T W A) / sigma_y^2 + 1/(2 sigma_x^2)) (each of
the six neighbour terms adds b / (2 sigma_x^2) of curvature, present or
clipped), alpha the minimiser of the exact data quadratic plus half the prior's
diagonal curvature along delta, clipped to [eps32, 1.5] (the over-relaxation
factor 2 of tomography_model.py:1440-1441; this part is pinned only by
restatement).  The last test plants each restatement mistake of
oracle.vcd.PLANTED and shows that one of these checks fails.
"""
import numpy as np
import pytest
from scipy.optimize import nnls

from oracle import partitions as opart
from oracle import projectors as op
from oracle import vcd as ovcd

EPS32 = float(np.finfo(np.float32).eps)
NBR_WTS = (1.0, 0.6, 1.4)
SIGMA_Y, SIGMA_X = 0.01, 0.005


class Problem:
    def __init__(self, sparse_truth=False, weighted=False, seed=0):
        self.geom = op.Geometry("cone", (12, 6, 10), np.linspace(0, 2 * np.pi, 12, endpoint=False), 40.0, 20.0)
        rows, cols, nz = self.shape = self.geom.recon_shape
        self.n = rows * cols * nz
        self.ror = np.flatnonzero(opart.get_2d_ror_mask(self.shape))
        self.ror_cols = (self.ror[:, None] * nz + np.arange(nz)).ravel()
        self.A = op.dense_matrix(self.geom)
        rng = np.random.default_rng(seed)
        xt = rng.random(self.shape) * opart.get_2d_ror_mask(self.shape)[:, :, None]
        if sparse_truth:
            xt *= rng.random(self.shape) > 0.6
        y = self.A @ xt.reshape(-1)
        y += 0.02 * y.max() * rng.standard_normal(y.shape)
        self.y = y
        self.w = rng.uniform(0.5, 1.5, y.shape) if weighted else np.ones(y.shape)
        self.L = self.laplacian()

    def laplacian(self):
        rows, cols, nz = self.shape
        wts = np.array(NBR_WTS)
        b_axis = wts / (2 * wts.sum())  # each direction's share of the six normalised weights
        idx = np.arange(self.n).reshape(self.shape)
        L = np.zeros((self.n, self.n))
        for axis in range(3):
            w = b_axis[axis] / (2 * SIGMA_X ** 2)
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[axis], hi[axis] = slice(0, -1), slice(1, None)
            for u, v in zip(idx[tuple(lo)].ravel(), idx[tuple(hi)].ravel()):
                L[u, u] += w
                L[v, v] += w
                L[u, v] -= w
                L[v, u] -= w
        return L

    def normal_equations(self):
        c = self.ror_cols
        H = self.A.T @ (self.w[:, None] * self.A) / SIGMA_Y ** 2 + self.L
        rhs = self.A.T @ (self.w * self.y) / SIGMA_Y ** 2
        return H[np.ix_(c, c)], rhs[c]

    def map_estimate(self, positivity=False):
        H, rhs = self.normal_equations()
        x = np.zeros(self.n)
        if positivity:
            # 1/2 x^T H x - rhs^T x = 1/2 |R x - R^-T rhs|^2 + const, H = R^T R
            R = np.linalg.cholesky(H).T
            x[self.ror_cols] = nnls(R, np.linalg.solve(R.T, rhs), maxiter=50 * rhs.size)[0]
        else:
            x[self.ror_cols] = np.linalg.solve(H, rhs)
        return x

    def oracle(self, positivity=False, planted=None, granularity=(1, 4, 16)):
        om = ovcd.OracleModel(self.geom, p=2.0, q=2.0, sigma_y=SIGMA_Y, sigma_x=SIGMA_X, auto_regularize_flag=False,
                              granularity=list(granularity), partition_sequence=[0, 1, 2],
                              qggmrf_nbr_wts=list(NBR_WTS), positivity_flag=positivity)
        om.planted = planted
        return om

    def run_oracle(self, iterations, positivity=False, planted=None):
        om = self.oracle(positivity, planted)
        np.random.seed(0)
        w = None if np.all(self.w == 1) else self.w.reshape(self.geom.sinogram_shape)
        x, stats = om.recon(self.y.reshape(self.geom.sinogram_shape), weights=w, init_recon=0,
                            max_iterations=iterations, stop_threshold_change_pct=0)
        return x.reshape(-1), stats


def fixed_point_error(prob, positivity=False, planted=None, iterations=60):
    x, _ = prob.run_oracle(iterations, positivity, planted)
    xs = prob.map_estimate(positivity)
    return np.linalg.norm(x - xs) / np.linalg.norm(xs)


@pytest.mark.parametrize("weighted", [False, True])
def test_vcd_reaches_quadratic_map(weighted):
    """60 iterations of the default-style schedule (K = 1, 4, then 16) reach
    the minimiser of the quadratic MAP cost; the absolute eps32 in the step's
    denominator stalls the last digits near 5e-7."""
    assert fixed_point_error(Problem(weighted=weighted)) < 2e-6


def test_vcd_positivity_reaches_constrained_map():
    """With positivity the limit is the non-negative minimiser (KKT point) of
    the same cost; the truth has 60 % zeros, so the constraint is active."""
    prob = Problem(sparse_truth=True)
    unconstrained = prob.map_estimate()
    assert np.sum(unconstrained < 0) >= 10  # the bound matters at the minimiser
    assert fixed_point_error(prob, positivity=True) < 2e-6


def dense_iteration(prob, x0, part, order, positivity):
    """One VCD iteration (every subset of `part` in `order`) on the dense problem,
    starting from init_recon x0 with the initial scaling of vcd_recon."""
    A, L, w, y = prob.A, prob.L, prob.w, prob.y
    nz = prob.shape[2]
    e = A @ x0
    a0 = np.sum(w * e * y) / np.sum(w * e * e)
    e = y - a0 * e
    x = a0 * x0
    fm_h = (A * A).T @ w / SIGMA_Y ** 2
    pr_h = 1.0 / (2 * SIGMA_X ** 2)
    alphas = []
    for si in order:
        c = (part[si][:, None] * nz + np.arange(nz)).ravel()
        Lx = L[c] @ x
        grad = -A[:, c].T @ (w * e) / SIGMA_Y ** 2 + Lx
        delta = -grad / (fm_h[c] + pr_h)
        ds = A[:, c] @ delta
        num = np.sum(w * e * ds) / SIGMA_Y ** 2 - Lx @ delta
        den = np.sum(w * ds * ds) / SIGMA_Y ** 2 + 0.5 * pr_h * np.sum(delta ** 2) + EPS32
        alpha = float(np.clip(num / den, EPS32, 1.5))
        if positivity:
            delta = np.maximum(-x[c] / (alpha + EPS32), delta)
            ds = A[:, c] @ delta
        x[c] += alpha * delta
        e = e - alpha * ds
        alphas.append(alpha)
    return x, e, np.mean(alphas)


def replay_error(prob, positivity=False, planted=None):
    """Relative gap between one oracle iteration (K = 4) and its dense replay."""
    rng = np.random.default_rng(7)
    x0 = np.zeros(prob.n)
    x0[prob.ror_cols] = rng.random(prob.ror_cols.size) - (0.3 if positivity else 0.0)
    x0 = np.maximum(x0, 0) if positivity else x0
    om = prob.oracle(positivity, planted, granularity=(4,))
    np.random.seed(3)
    part = opart.gen_set_of_pixel_partitions(prob.shape, [4])[0]
    state = np.random.get_state()
    # the oracle draws gen_full_indices (forward projection of the initial
    # recon) and then the subset order from the global RNG
    opart.gen_full_indices(prob.shape)
    order = np.random.permutation(part.shape[0])
    np.random.set_state(state)
    w = None if np.all(prob.w == 1) else prob.w.reshape(prob.geom.sinogram_shape)
    x, stats = om.vcd_recon(prob.y.reshape(prob.geom.sinogram_shape), [part], [0], 0.0, weights=w,
                            init_recon=x0.reshape(prob.shape))
    xd, ed, ad = dense_iteration(prob, x0, part, order, positivity)
    gx = np.linalg.norm(x.reshape(-1) - xd) / np.linalg.norm(xd)
    ge = np.linalg.norm(stats["error_sinogram"].reshape(-1) - ed) / np.linalg.norm(ed)
    return max(gx, ge, abs(stats["alpha_values"][0] - ad) / ad)


@pytest.mark.parametrize("weighted,positivity", [(False, False), (True, False), (False, True)])
def test_vcd_iteration_matches_dense_replay(weighted, positivity):
    prob = Problem(weighted=weighted, sparse_truth=positivity)
    assert replay_error(prob, positivity) < 1e-10


@pytest.mark.parametrize("planted", sorted(ovcd.PLANTED))
def test_planted_vcd_mistake_is_caught(planted):
    weighted = planted == "hessian_unweighted"
    positivity = planted == "positivity_no_alpha"
    prob = Problem(weighted=weighted, sparse_truth=positivity)
    gaps = dict(replay=replay_error(prob, positivity, planted))
    if planted in ("prior_sign", "fm_constant_sigma", "error_update_sign"):
        # mistakes that move the fixed point itself
        with np.errstate(all="ignore"):
            gaps["fixed_point"] = fixed_point_error(prob, positivity, planted, iterations=30)
        assert not gaps["fixed_point"] < 1e-3, (ovcd.PLANTED[planted], gaps)
    assert not gaps["replay"] < 1e-6, (ovcd.PLANTED[planted], gaps)


def test_duplicate_pixels_add_every_copy():
    """A padded subset can list a pixel twice (mbirjax/vcd_utils.py:212-217);
    the update then adds every copy's step, as .at[].add does
    (mbirjax/tomography_model.py:1820-1823).  The error sinogram is updated
    with the projection of every listed copy, so the invariant e = y - A x
    holds only if the recon takes every copy too."""
    prob = Problem()
    om = prob.oracle()
    nz = prob.shape[2]
    y = prob.y.reshape(prob.geom.sinogram_shape)
    om.y = y
    om.check_invariant = True
    flat = np.zeros((prob.shape[0] * prob.shape[1], nz))
    fm_hess = om.hessian_diagonal().reshape(-1, nz)
    ror = prob.ror
    S = np.array([ror[3], ror[10], ror[3], ror[17], ror[10], ror[3]])
    flat, e, ell1, alpha = om.subset_update(flat, y.copy(), S, 1.0, True, fm_hess, 1.0 / SIGMA_Y ** 2,
                                            om.qparams(), None, 1.0)
    resid = y.reshape(-1) - prob.A @ flat.reshape(-1)
    assert np.max(np.abs(resid - e.reshape(-1))) <= 1e-9 * np.max(np.abs(y))
    assert np.any(flat[ror[3]] != 0)
