/*
 * mbx.h -- C ABI of libmbx, the MI355X (gfx950) kernel library behind the
 * mbirjax_amd TomographyModel API.
 *
 * Every entry point is stream-ordered on the caller's hipStream_t (passed as
 * void*; NULL means the default stream), allocates no device memory unless it
 * says so, validates its scalar arguments before any HIP call, and returns an
 * int status: 0 = OK, negative = error (see MBX_E*).  mbx_last_error() returns
 * a thread-local message describing the most recent failure.
 *
 * Array conventions (all device pointers):
 *   - pixel indices: int32, flat index p = i*num_recon_cols + j into the 2D
 *     (rows, cols) recon grid (C order), every entry in [0, rows*cols).
 *   - voxel cylinders: float32 (npix, num_recon_slices), row-major.
 *   - sinograms: views x channels x rows ("channel-major", MBX_LAYOUT_VCR) or
 *     the user-facing views x rows x channels (MBX_LAYOUT_VRC); float32 or
 *     bfloat16 (MBX_DTYPE_*), and float16 where MBX_DTYPE_F16 says so.  The projectors read/write MBX_LAYOUT_VCR;
 *     mbx_sino_transpose converts.
 *   - angles: float32 (num_views,), radians.
 *
 * Reference interfaces replaced (mbirjax 0.6.7, paths relative to its root):
 *   mbx_forward_project  <- Projectors.sparse_forward_project
 *                           (mbirjax/projectors.py:61-129) over the geometry
 *                           plugins ConeBeamModel.forward_project_pixel_batch_to_one_view
 *                           (mbirjax/cone_beam.py:189-377) and
 *                           ParallelBeamModel.forward_project_pixel_batch_to_one_view
 *                           (mbirjax/parallel_beam.py:138-178)
 *   mbx_back_project     <- Projectors.sparse_back_project
 *                           (mbirjax/projectors.py:131-209) over
 *                           ConeBeamModel.back_project_one_view_to_pixel_batch
 *                           (mbirjax/cone_beam.py:379-541) and
 *                           ParallelBeamModel.back_project_one_view_to_pixel_batch
 *                           (mbirjax/parallel_beam.py:180-221); coeff_power=2 is
 *                           TomographyModel.compute_hessian_diagonal
 *                           (mbirjax/tomography_model.py:679-708)
 *   mbx_qggmrf_grad_hess <- qggmrf_gradient_and_hessian_at_indices
 *                           (mbirjax/qggmrf.py:118-279)
 *   mbx_vcd_subset_update<- TomographyModel.create_vcd_subset_updater /
 *                           vcd_subset_updater (mbirjax/tomography_model.py:1310-1544)
 *                           with get_forward_lin_quad (1546-1631) and
 *                           update_recon (1820-1823)
 *   mbx_denoise_subset_update <- vcd_subset_denoiser (mbirjax/denoising.py:366-411),
 *                           driven by QGGMRFDenoiser.denoise (142-364)
 *   mbx_median_filter3d  <- median_filter3d (mbirjax/denoising.py:414-496)
 *   mbx_dot              <- the jnp.sum reductions of vcd_recon / get_forward_model_loss
 *                           (mbirjax/tomography_model.py:1139-1152, 1633-1658)
 *   mbx_forward_plan / mbx_forward_apply: no counterpart (mbx_forward_project
 *                           split into its setup and its projection kernel)
 *   mbx_vcd_subset_update_planned: vcd_subset_updater with the subset's
 *                           projector setup (geometry table, bins, records) built
 *                           once by mbx_forward_plan and reused
 *
 * Python binds these through ctypes (mbirjax_amd/_lib.py).
 */
#ifndef MBX_H
#define MBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBX_OK 0
#define MBX_EINVAL (-1)       /* invalid argument (checked before any HIP call) */
#define MBX_ELAUNCH (-2)      /* a HIP launch or runtime call failed */
#define MBX_EUNSUPPORTED (-3) /* valid but unsupported combination */
#define MBX_EWORKSPACE (-4)   /* workspace missing or too small */

#define MBX_GEOM_PARALLEL 0
#define MBX_GEOM_CONE 1
#define MBX_GEOM_TRANSLATION 2 /* mbirjax/translation_model.py: no rotation, a translation per view */

#define MBX_DTYPE_F32 0
#define MBX_DTYPE_BF16 1
#define MBX_DTYPE_F64 2 /* mbx_median_filter3d only */
#define MBX_DTYPE_F16 3 /* IEEE binary16: a 2-byte error sinogram (read by the back projector, mbx_dot,
                           mbx_sino_axpy, mbx_sino_transpose and the VCD update; never a forward output).
                           The caller keeps |values| below 65504, e.g. by scaling the problem by a power
                           of two, which leaves the VCD trajectory unchanged apart from the rounding */

#define MBX_LAYOUT_VCR 0 /* (views, channels, rows): projector-native */
#define MBX_LAYOUT_VRC 1 /* (views, rows, channels): mbirjax user layout */

/* Geometry of one TomographyModel, mirroring the reference's projector_params
 * = (sinogram_shape, recon_shape, geometry_params): ConeBeamModel.get_geometry_parameters
 * (mbirjax/cone_beam.py:110-136), ParallelBeamModel.get_geometry_parameters
 * (mbirjax/parallel_beam.py:94-113) and TranslationModel.get_geometry_parameters
 * (mbirjax/translation_model.py:83-107).  Doubles are rounded to float32 on the
 * device, as the reference's jitted float32 arithmetic does.
 *
 * View parameters ("angles" in the entry points below): num_views float32
 * angles for MBX_GEOM_PARALLEL and MBX_GEOM_CONE; num_views x 3 float32
 * translation vectors (x, y, z) for MBX_GEOM_TRANSLATION, whose forward
 * vertical fan uses psf_radius (bp_psf_radius is ignored). */
typedef struct mbx_geom {
    int32_t kind;             /* MBX_GEOM_PARALLEL, MBX_GEOM_CONE or MBX_GEOM_TRANSLATION */
    int32_t num_views;        /* length of the angles array */
    int32_t num_det_rows;     /* R */
    int32_t num_det_channels; /* C */
    int32_t num_recon_rows;
    int32_t num_recon_cols;
    int32_t num_recon_slices; /* Nz */
    int32_t psf_radius;       /* channel (and cone back-projection row) tap radius */
    int32_t bp_psf_radius;    /* cone forward-projection slice tap radius */
    int32_t reserved;
    double delta_det_row;
    double delta_det_channel;
    double det_row_offset;
    double det_channel_offset;
    double source_detector_dist; /* may be +inf (cone) ; unused for parallel */
    double magnification;        /* source_detector_dist / source_iso_dist; 1 for parallel */
    double delta_voxel;
    double recon_slice_offset;
    double delta_recon_row;      /* MBX_GEOM_TRANSLATION: pitch of the recon rows (y); unused otherwise */
} mbx_geom;

/* qGGMRF prior parameters: b = get_b_from_nbr_wts(qggmrf_nbr_wts) in the order
 * [row+1, row-1, col+1, col-1, slice+1, slice-1] (mbirjax/qggmrf.py:388-395). */
typedef struct mbx_qggmrf {
    double b[6];
    double sigma_x;
    double p;
    double q;
    double T;
} mbx_qggmrf;

int mbx_version(void);
const char *mbx_last_error(void);

/* Number of device bytes mbx_back_project needs as workspace (0 = none). */
int64_t mbx_back_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t nviews);

/* Sparse forward projection (mbirjax/projectors.py:61-129).
 * vox: (npix, Nz) f32; pix: (npix,) int32.  views: NULL = all num_views views,
 * else nviews int32 indices into angles; output view slot i uses angle views[i].
 * sino: (nviews_out, C, R) in sino_dtype, layout sino_layout.
 * accumulate = 0 writes the projection, 1 adds it to sino.
 * sumsq (nullable, device double[1]): receives sum over the written
 * projection values of value^2 (times weights[...] when weights != NULL; weights
 * then has sino's shape, f32, layout sino_layout); this is forward_quadratic
 * of get_forward_lin_quad before the 1/sigma_y^2 factor.  sumsq needs a
 * workspace of mbx_forward_workspace_bytes() bytes (deterministic per-block
 * partials); with sumsq == NULL the workspace may be NULL. */
int64_t mbx_forward_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t nviews);
int mbx_forward_project(const mbx_geom *g, const float *angles, const float *vox, const int32_t *pix,
                        int32_t npix, const int32_t *views, int32_t nviews, void *sino, int32_t sino_dtype,
                        int32_t sino_layout, int32_t accumulate, const float *weights, double *sumsq,
                        void *workspace, int64_t workspace_bytes, void *stream);

/* The forward projection in two steps, for repeated projections of one pixel
 * set (no counterpart in mbirjax, whose projector recomputes its geometry on
 * every call; for callers that apply A to the same subset many times, such as
 * unrolled or autograd loops, and for timing the projection kernel alone).
 * mbx_forward_plan builds the geometry table, the centre-channel bins and the
 * forward records of (pix, views) into plan, which holds
 * mbx_forward_workspace_bytes(g, npix, nviews) bytes; the pixel set must fit
 * one geometry-table chunk (about 512 MB of table: 10965 pixels at 720 views),
 * else MBX_EUNSUPPORTED.  mbx_forward_apply then projects cylinders vox of
 * those pixels exactly as mbx_forward_project would (same arguments; views
 * only fixes the view count); plan_bytes must equal
 * mbx_forward_workspace_bytes(g, npix, nviews) for the npix and nviews the
 * plan was built with, else MBX_EINVAL (so an apply of more pixels or views
 * than planned cannot read past the plan).  The plan stays valid while the
 * geometry, angles, pix and views are unchanged; applies on one stream may
 * share it (an apply that asks for sumsq writes the per-wave FP64 partials
 * into the plan's scratch slots, so applies must not overlap in time). */
int mbx_forward_plan(const mbx_geom *g, const float *angles, const int32_t *pix, int32_t npix, const int32_t *views,
                     int32_t nviews, void *plan, int64_t plan_bytes, void *stream);
int mbx_forward_apply(const mbx_geom *g, const void *plan, int64_t plan_bytes, const float *vox, int32_t npix,
                      const int32_t *views, int32_t nviews, void *sino, int32_t sino_dtype, int32_t sino_layout,
                      int32_t accumulate, const float *weights, double *sumsq, void *stream);

/* Sparse back projection (mbirjax/projectors.py:131-209, tomography_model.py:624-677).
 * sino: (num_sino_views, C, R) in sino_dtype, MBX_LAYOUT_VCR.  views: NULL =
 * views 0..num_sino_views-1 (then num_sino_views must equal g->num_views), else
 * nviews indices that select both the sinogram view and its angle (the
 * reference's view_indices semantics, tomography_model.py:661-663).
 * weights (nullable) multiplies the sinogram element-wise (same shape, f32).
 * coeff_power 1 or 2 (2 = Hessian diagonal; float32 sinogram).  sino_dtype
 * MBX_DTYPE_F32, MBX_DTYPE_BF16 or MBX_DTYPE_F16.  out (npix, Nz) f32 = (accumulate ? out : 0) +
 * scale * A^T(weights*sino).  workspace: mbx_back_workspace_bytes() bytes. */
int mbx_back_project(const mbx_geom *g, const float *angles, const void *sino, int32_t sino_dtype,
                     int32_t num_sino_views, const float *weights, const int32_t *pix, int32_t npix,
                     const int32_t *views, int32_t nviews, int32_t coeff_power, float scale, int32_t accumulate,
                     float *out, void *workspace, int64_t workspace_bytes, void *stream);

/* The back projection of a planned pixel set (no counterpart in mbirjax, whose
 * projector recomputes its geometry on every call; for callers that
 * back-project onto the same pixel set many times, such as the backward pass
 * of an unrolled or autograd loop over a fixed subset, and for timing the
 * back-projection kernel alone).  mbx_back_project of the npix pixels and the
 * views a plan of mbx_forward_plan was built with, reading the plan's
 * geometry table instead of building one: same arguments and result, with
 * plan and plan_bytes as for mbx_forward_apply (plan_bytes must equal
 * mbx_forward_workspace_bytes(g, npix, nviews), else MBX_EINVAL) and a
 * workspace of mbx_back_apply_workspace_bytes(g, npix, nviews) bytes (the
 * view-group partials; 0 bytes needs no workspace).  The plan is only read,
 * so back and forward applies of one plan may follow each other on a stream. */
int64_t mbx_back_apply_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t nviews);
int mbx_back_apply(const mbx_geom *g, const void *plan, int64_t plan_bytes, const void *sino, int32_t sino_dtype,
                   int32_t num_sino_views, const float *weights, int32_t npix, const int32_t *views, int32_t nviews,
                   int32_t coeff_power, float scale, int32_t accumulate, float *out, void *workspace,
                   int64_t workspace_bytes, void *stream);

/* mbx_back_project without a workspace (no counterpart in mbirjax: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and result
 * as mbx_back_project without workspace:
 * the geometry table and the view-group partials live in a file-scope device
 * array of mbx_noalloc_scratch_bytes() bytes (64 MiB), which every GPU the
 * library runs on holds until the process ends.  Large pixel sets are
 * projected in table chunks that fit it (bitwise the sums of
 * mbx_back_project whenever its own chunk fits); a call whose single-pixel
 * table and partials do not fit returns MBX_EUNSUPPORTED.  The call allocates
 * nothing (no hipMalloc, hipFree or allocating library routine) and contains
 * no memset, so it can be captured into a HIP graph.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race and must be serialised by the caller. */
int64_t mbx_noalloc_scratch_bytes(void);
int mbx_back_project_noalloc(const mbx_geom *g, const float *angles, const void *sino, int32_t sino_dtype,
                             int32_t num_sino_views, const float *weights, const int32_t *pix, int32_t npix,
                             const int32_t *views, int32_t nviews, int32_t coeff_power, float scale,
                             int32_t accumulate, float *out, void *stream);

/* Convert a sinogram between layouts (and dtypes).  src and dst must not alias. */
int mbx_sino_transpose(const void *src, int32_t src_dtype, int32_t src_layout, void *dst, int32_t dst_dtype,
                       int32_t dst_layout, int32_t num_views, int32_t num_rows, int32_t num_channels,
                       void *stream);

/* qGGMRF surrogate gradient and Hessian diagonal at the cylinders pix
 * (mbirjax/qggmrf.py:118-279): grad, hess (npix, Nz) f32 are overwritten. */
int mbx_qggmrf_grad_hess(const float *flat_recon, int32_t num_rows, int32_t num_cols, int32_t num_slices,
                         const int32_t *pix, int32_t npix, const mbx_qggmrf *params, float *grad, float *hess,
                         void *stream);

/* Deterministic FP64 reduction: out[0] = sum_i a[i]*b[i]*(w ? w[i] : 1),
 * with b == NULL meaning b = a.  a, b in their dtypes, w f32.  n elements.
 * The workspace holds per-block partials (mbx_dot_workspace_bytes()). */
int64_t mbx_dot_workspace_bytes(void);
int mbx_dot(const void *a, int32_t a_dtype, const void *b, int32_t b_dtype, const float *w, int64_t n,
            double *out, void *workspace, int64_t workspace_bytes, void *stream);

/* e = e - alpha[0] * d  (alpha a device double), elementwise over n entries of
 * the error sinogram e (e_dtype) and the delta sinogram d (d_dtype). */
int mbx_sino_axpy(void *e, int32_t e_dtype, const void *d, int32_t d_dtype, const double *alpha, int64_t n,
                  void *stream);

/* One VCD subset update: vcd_subset_updater (mbirjax/tomography_model.py:1343-1542).
 * State: flat_recon (rows*cols, Nz) f32 and the error sinogram e (V, C, R) in
 * e_dtype (MBX_DTYPE_F32, _BF16 or _F16; MBX_LAYOUT_VCR), e = y - A x.
 * weights: NULL = constant 1, else (V, C, R) f32.  fm_hessian: (rows*cols, Nz)
 * f32 = diag(A^T W A).  prox_input: NULL = qGGMRF prior, else (rows*cols, Nz)
 * f32 proximal-map input with prior Hessian sigma_prox^2.  delta_dtype:
 * storage of the delta sinogram (MBX_DTYPE_F32 or _BF16).
 * stats (device double[4]) accumulates {sum |alpha*delta|, alpha, 1, 0} for
 * per-iteration statistics.  workspace: mbx_vcd_workspace_bytes() bytes.
 * Subsets of 2^18 (view, pixel) pairs or more bin the forward records on a
 * second stream of the device that runs `stream` (the current device for the
 * NULL stream; default priority, created on first use with two events and
 * kept until the process ends), forked from and joined back into `stream`
 * inside the call: the call stays stream-ordered, also on an error return.
 * The workspace size grows with npix: a workspace of
 * mbx_vcd_workspace_bytes(g, N, dt) bytes serves every subset of up to N
 * pixels. */
int64_t mbx_vcd_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t delta_dtype);
int mbx_vcd_subset_update(const mbx_geom *g, const float *angles, float *flat_recon, void *e, int32_t e_dtype,
                          const float *weights, const float *fm_hessian, const int32_t *pix, int32_t npix,
                          double fm_constant, const mbx_qggmrf *qparams, const float *prox_input,
                          double sigma_prox, int32_t positivity, int32_t delta_dtype,
                          double *stats, void *workspace, int64_t workspace_bytes, void *stream);

/* The same subset update with the subset's geometry table, centre-channel bins
 * and forward records taken from plan: the output of
 * mbx_forward_plan(g, angles, pix, npix, NULL, 0, plan, plan_bytes, stream)
 * for the same pix (all views), which must fit one geometry-table chunk
 * (else MBX_EUNSUPPORTED); plan_bytes must equal
 * mbx_forward_workspace_bytes(g, npix, g->num_views) (else MBX_EINVAL).
 * For callers that sweep a fixed partition many times (VCD iterations with
 * one granularity): the setup is then paid once per subset instead of once
 * per update.  Results are bitwise those of mbx_vcd_subset_update.  The call
 * writes the forward projection's FP64 partials into the plan's scratch slots:
 * calls sharing a plan must not overlap in time. */
int mbx_vcd_subset_update_planned(const mbx_geom *g, const float *angles, float *flat_recon, void *e,
                                  int32_t e_dtype, const float *weights, const float *fm_hessian,
                                  const int32_t *pix, int32_t npix, double fm_constant, const mbx_qggmrf *qparams,
                                  const float *prox_input, double sigma_prox, int32_t positivity,
                                  int32_t delta_dtype, void *plan, int64_t plan_bytes, double *stats,
                                  void *workspace, int64_t workspace_bytes, void *stream);

/* One subset update of the qGGMRF proximal-map denoiser (vcd_subset_denoiser,
 * mbirjax/denoising.py:366-411): the identity forward model on the image,
 * so flat_error = noisy image - flat_image and the forward Hessian is 1.
 * flat_image, flat_error: (num_rows*num_cols, num_slices) f32, updated in
 * place at the cylinders pix.  fm_constant = 1/sigma_noise^2.  stats (device
 * double[4], nullable) accumulates {sum |alpha*delta|, alpha, 1, 0}.
 * workspace: mbx_denoise_workspace_bytes() bytes. */
int64_t mbx_denoise_workspace_bytes(int32_t npix, int32_t num_slices);
int mbx_denoise_subset_update(float *flat_image, float *flat_error, int32_t num_rows, int32_t num_cols,
                              int32_t num_slices, const int32_t *pix, int32_t npix, double fm_constant,
                              const mbx_qggmrf *qparams, double *stats, void *workspace, int64_t workspace_bytes,
                              void *stream);

/* 27-point (3x3x3) median filter with edge-replicated borders
 * (median_filter3d, mbirjax/denoising.py:414-496).  x, out: (d0, d1, d2) in
 * dtype MBX_DTYPE_F32 or MBX_DTYPE_F64, distinct buffers. */
int mbx_median_filter3d(const void *x, int32_t dtype, int32_t d0, int32_t d1, int32_t d2, void *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MBX_H */
