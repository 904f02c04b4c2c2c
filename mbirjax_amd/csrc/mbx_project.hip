// Argument validation, the per-(view, pixel) geometry table, the sinogram
// layout transpose and the C ABI of the sparse projectors on gfx950.
//
// Operator (identical in both directions, so back == forward^T exactly):
//   A[v, m, n ; p, k] = Ac(p,v,n) * Lr(t_{p,v,k} - m) / cos(phi_{p,v,k})
// with the horizontal fan weight Ac (mbirjax/cone_beam.py:244-284, 409-451,
// 593-644), the vertical fan weight Lr over the projected slice position
// t_k = a + W*k (mbirjax/cone_beam.py:286-377, 480-591) and, for parallel beam,
// Lr = [m == k], cos(phi) = 1 (mbirjax/parallel_beam.py:138-263).
// The back projector is in mbx_back.hip, the forward projector in mbx_fwd.hip;
// both read the table built here, so they apply exactly transposed operators.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstdio>
#include <string>

#include "mbx_internal.hpp"

namespace mbx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int make_dev_geom(const mbx_geom *g, DevGeom *d) {
    if (!g) return fail(MBX_EINVAL, "geometry pointer is NULL");
    if (g->kind != MBX_GEOM_PARALLEL && g->kind != MBX_GEOM_CONE && g->kind != MBX_GEOM_TRANSLATION)
        return fail(MBX_EINVAL, "geometry kind must be MBX_GEOM_PARALLEL, MBX_GEOM_CONE or MBX_GEOM_TRANSLATION");
    const bool fan = g->kind != MBX_GEOM_PARALLEL;
    if (g->num_views <= 0 || g->num_det_rows <= 0 || g->num_det_channels <= 0 || g->num_recon_rows <= 0 ||
        g->num_recon_cols <= 0 || g->num_recon_slices <= 0)
        return fail(MBX_EINVAL, "all sinogram and recon dimensions must be positive");
    if ((int64_t)g->num_recon_rows * g->num_recon_cols > INT32_MAX)
        return fail(MBX_EINVAL, "num_recon_rows*num_recon_cols exceeds int32");
    // The kernels address within one view with 32-bit byte offsets; the table
    // packs two row indices into one word.
    if ((int64_t)g->num_det_rows * g->num_det_channels * 4 > INT32_MAX)
        return fail(MBX_EINVAL, "num_det_rows*num_det_channels*4 bytes per view exceeds int32");
    if (g->num_det_rows > 32767) return fail(MBX_EINVAL, "num_det_rows must be below 32768");
    if (!(g->delta_det_channel > 0) || !(g->delta_voxel > 0) || !std::isfinite(g->delta_det_channel) ||
        !std::isfinite(g->delta_voxel))
        return fail(MBX_EINVAL, "delta_det_channel and delta_voxel must be positive and finite");
    if (g->psf_radius < 1 || g->psf_radius > 2)
        return fail(MBX_EUNSUPPORTED, "psf_radius must be 1 or 2 (got " + std::to_string(g->psf_radius) + ")");
    if (fan) {
        if (!(g->delta_det_row > 0) || !std::isfinite(g->delta_det_row))
            return fail(MBX_EINVAL, "delta_det_row must be positive and finite");
        if (!(g->magnification > 0) || !std::isfinite(g->magnification))
            return fail(MBX_EINVAL, "magnification must be positive and finite");
        if (!(g->source_detector_dist > 0)) return fail(MBX_EINVAL, "source_detector_dist must be positive");
        if (g->kind == MBX_GEOM_CONE && (g->bp_psf_radius < 1 || g->bp_psf_radius > 2))
            return fail(MBX_EUNSUPPORTED,
                        "bp_psf_radius must be 1 or 2 (got " + std::to_string(g->bp_psf_radius) + ")");
        // the translation geometry has every pixel at its own distance y from
        // the source, all views sharing it (mbirjax/translation_model.py:69-81)
        if (g->kind == MBX_GEOM_TRANSLATION &&
            (!std::isfinite(g->source_detector_dist) || !(g->delta_recon_row > 0) || !std::isfinite(g->delta_recon_row)))
            return fail(MBX_EINVAL, "translation geometry needs a finite source_detector_dist and a positive "
                                    "delta_recon_row");
    } else {
        if (g->num_recon_slices != g->num_det_rows)
            return fail(MBX_EINVAL, "parallel beam needs num_recon_slices == num_det_rows");
    }
    DevGeom &o = *d;
    o.kind = g->kind;
    o.V = g->num_views;
    o.R = g->num_det_rows;
    o.C = g->num_det_channels;
    o.rows = g->num_recon_rows;
    o.cols = g->num_recon_cols;
    o.Nz = g->num_recon_slices;
    o.rh = g->psf_radius;
    // forward vertical taps: bp_psf_radius (cone), psf_radius (translation,
    // mbirjax/translation_model.py:364)
    o.rb = g->kind == MBX_GEOM_CONE ? g->bp_psf_radius : g->kind == MBX_GEOM_TRANSLATION ? g->psf_radius : 1;
    o.dr = (float)(fan ? g->delta_det_row : 1.0);
    o.dc = (float)g->delta_det_channel;
    o.row_off = (float)g->det_row_offset;
    o.ch_off = (float)g->det_channel_offset;
    o.inv_sdd = fan ? (float)(1.0 / g->source_detector_dist) : 0.0f;
    o.inv_M = fan ? (float)(1.0 / g->magnification) : 1.0f;
    o.dv = (float)g->delta_voxel;
    o.slice_off = (float)g->recon_slice_offset;
    o.row_center = (float)((g->num_det_rows - 1) / 2.0);
    o.ch_center = (float)((g->num_det_channels - 1) / 2.0);
    o.img_row_center = (float)((g->num_recon_rows - 1) / 2.0);
    o.img_col_center = (float)((g->num_recon_cols - 1) / 2.0);
    o.slice_center = (float)((g->num_recon_slices - 1) / 2.0);
    o.dv_over_dc = (float)(g->delta_voxel / g->delta_det_channel);
    o.dv_over_dr = fan ? (float)(g->delta_voxel / g->delta_det_row) : 1.0f;
    o.kref = g->num_recon_slices / 2;
    o.h = *g;
    return MBX_OK;
}

// ======================================================================
// Per-(view, pixel) geometry table
// ======================================================================
//
// One 16-float record per (view slot, subset pixel), computed once per call
// (or once per VCD subset update) in FP64 from the caller's double geometry
// and read by both projectors with scalar loads, so both directions use
// bit-identical weights:
//   [0] n_c = rint(n_p) (int bits)    [1..5] Ac(n_c-2 .. n_c+2), zero outside
//   the tap radius and outside [0, C); at radius 1 (cone and translation)
//   [1] and [5] hold instead the first and last rows (int bits) that the
//   back projection of all Nz slices reaches, taps included
//   vertical fan, relative to the reference slice kref = Nz/2 and the integer
//   row ai = rint(t_kref):  t_k - ai = acr + (W + W_lo) * (k - kref)
//   [6] acr   [7] W (rounded)   [8] W_lo = W - float(W)   [9] (W+1)/2
//   [10] min(1, W)   [11] 1/W   [12] u at kref (v/sdd)   [13] du/dk
//   [14] ai (int bits)   [15] first | last row touched by any slice << 16
// plus n_c again in a dense (view, pixel) int array for the forward scan.
// Keeping the row positions relative to ai lets the kernels form t_k - m
// without rounding at the magnitude of a detector row index.


int table_chunk(const DevGeom &g, int npix, int nviews) {
    const int64_t per_pixel = (int64_t)std::max(nviews, 1) * (kRec * 4 + 4);
    const int64_t c = std::max<int64_t>(1024, kTableBudget / per_pixel);
    return (int)std::min<int64_t>(npix, c);
}

int64_t table_bytes(int chunk, int nviews) {
    const int64_t rec = (int64_t)chunk * nviews * kRec * 4;
    const int64_t nc = (int64_t)chunk * nviews * 4;
    return ((rec + 255) & ~255ll) + ((nc + 255) & ~255ll);
}

template <int GEOM, int RAD>
__global__ __launch_bounds__(256) void pv_table_kernel(DevGeom g, const float *__restrict__ angles,
                                                       const int32_t *__restrict__ pix, int npix, int pstride,
                                                       const int32_t *__restrict__ views, float *__restrict__ T,
                                                       int *__restrict__ NC) {
    const int vi = blockIdx.y;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npix) return;
    const mbx_geom &h = g.h;
    const int va = views ? views[vi] : vi;
    // A pixel index outside the recon contributes nothing: its geometry is
    // that of pixel 0, its channel weights are zero and its centre channel is
    // far off the detector (the forward binning drops it), so no index the
    // caller passes reaches memory through the table.
    const int pin = pix[(size_t)s * pstride];
    const bool valid = pin >= 0 && pin < g.rows * g.cols;
    const int p = valid ? pin : 0;
    const int i = p / g.cols, j = p - (p / g.cols) * g.cols;
    const double dv = h.delta_voxel, dc = h.delta_det_channel;
    // Wc: voxel width on the detector in channels; sc: channel weight per unit
    // overlap; zoff: z offset of the slices
    double mag = 1.0, np, Wc, sc, zoff = h.recon_slice_offset;
    if (GEOM == MBX_GEOM_TRANSLATION) {
        // recon_ijk_to_xyz, geometry_xyz_to_uv_mag, compute_horizontal_data
        // (mbirjax/translation_model.py:595-672): no rotation, the object
        // shifted by -t; channel weight L / cos(theta_p)
        const float *tv = angles + 3 * (size_t)va;
        const double x = dv * ((double)j - 0.5 * (h.num_recon_cols - 1)) - (double)tv[0];
        const double y = h.delta_recon_row * ((double)i - 0.5 * (h.num_recon_rows - 1)) - (double)tv[1];
        zoff = -(double)tv[2];
        mag = 1.0 / (1.0 / h.magnification - y / h.source_detector_dist);
        const double u = mag * x;
        np = (u + h.det_channel_offset) / dc + 0.5 * (h.num_det_channels - 1);
        const double us = u / h.source_detector_dist;
        Wc = mag * (dv / dc);
        sc = sqrt(1.0 + us * us);  // 1 / cos(theta_p)
    } else {
        // recon_ijk_to_xyz and geometry_xyz_to_uv_mag (mbirjax/cone_beam.py:646-681),
        // compute_horizontal_data (593-644); parallel: parallel_beam.py:223-284
        double sn, cs;
        sincos((double)angles[va], &sn, &cs);
        const double yt = dv * ((double)i - 0.5 * (h.num_recon_rows - 1));
        const double xt = dv * ((double)j - 0.5 * (h.num_recon_cols - 1));
        const double x = cs * xt - sn * yt;
        double cosa, ctp = 1.0;
        if (GEOM == MBX_GEOM_CONE) {
            const double y = sn * xt + cs * yt;
            mag = 1.0 / (1.0 / h.magnification - y / h.source_detector_dist);
            const double u = mag * x;
            np = (u + h.det_channel_offset) / dc + 0.5 * (h.num_det_channels - 1);
            const double us = u / h.source_detector_dist;
            ctp = 1.0 / sqrt(1.0 + us * us);  // cos(theta_p)
            const double stp = us * ctp;      // sin(theta_p)
            cosa = fmax(fabs(cs * ctp + sn * stp), fabs(sn * ctp - cs * stp));
        } else {
            np = (x + h.det_channel_offset) / dc + 0.5 * (h.num_det_channels - 1);
            cosa = fmax(fabs(cs), fabs(sn));
        }
        Wc = (GEOM == MBX_GEOM_CONE ? mag * (cosa / ctp) : cosa) * (dv / dc);
        sc = dv / cosa;
    }
    const double hc = 0.5 * (Wc + 1.0), Lc = fmin(1.0, Wc);
    const int nc = valid ? (int)rint(np) : -(1 << 28);
    float r[kRec];
    r[0] = __int_as_float(nc);
#pragma unroll
    for (int t = -2; t <= 2; ++t) {
        const int n = nc + t;
        const double L = fmin(fmax(hc - fabs(np - (double)n), 0.0), Lc);
        r[3 + t] = (t >= -RAD && t <= RAD && n >= 0 && n < g.C) ? (float)(sc * L) : 0.0f;
    }
    int mlo = 0, mhi = g.R - 1, ai = g.kref;
    if (GEOM != MBX_GEOM_PARALLEL) {
        // compute_vertical_data_single_pixel (mbirjax/cone_beam.py:543-591,
        // translation_model.py:544-593): t_k = (mag * z_k + det_row_offset) / dr
        // + (R-1)/2, linear in k
        const double dr = h.delta_det_row;
        const double W = mag * dv / dr;
        const double zref = zoff + dv * ((double)g.kref - 0.5 * (h.num_recon_slices - 1));
        const double tref = (mag * zref + h.det_row_offset) / dr + 0.5 * (h.num_det_rows - 1);
        ai = (int)rint(tref);
        const float Wf = (float)W;
        r[6] = (float)(tref - ai);
        r[7] = Wf;
        r[8] = (float)(W - (double)Wf);
        r[9] = (float)(0.5 * (W + 1.0));
        r[10] = (float)fmin(1.0, W);
        r[11] = (float)(1.0 / W);
        r[12] = (float)(mag * zref / h.source_detector_dist);
        r[13] = (float)(mag * dv / h.source_detector_dist);
        const double t0 = tref - W * g.kref, t1 = tref + W * (g.Nz - 1 - g.kref);
        mlo = max(0, (int)rint(t0) - RAD);
        mhi = min(g.R - 1, (int)rint(t1) + RAD);
        if (RAD == 1) {
            // The rows the back projection of all Nz slices reaches, formed
            // exactly as the back projector forms them from this record (the
            // +-2 tap words are free at radius 1).
            const double Wd = (double)r[7] + (double)r[8], acd = (double)r[6];
            r[1] = __int_as_float(ai + (int)rint(fma(Wd, (double)(-g.kref), acd)) - RAD);
            r[5] = __int_as_float(ai + (int)rint(fma(Wd, (double)(g.Nz - 1 - g.kref), acd)) + RAD);
        }
    } else {
        r[6] = 0.0f;
        r[7] = 1.0f;
        r[8] = 0.0f;
        r[9] = 1.0f;
        r[10] = 1.0f;
        r[11] = 1.0f;
        r[12] = 0.0f;
        r[13] = 0.0f;
    }
    r[14] = __int_as_float(ai);
    r[15] = __int_as_float(mlo | (mhi << 16));
    float4 *dst = (float4 *)(T + ((size_t)vi * npix + s) * kRec);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    NC[(size_t)vi * npix + s] = nc;
}


int launch_table(const DevGeom &g, const float *angles, const int32_t *pix, int npix, const int32_t *views, int nviews,
                 float *T, int *NC, hipStream_t st, int pstride) {
    if (npix == 0 || nviews == 0) return MBX_OK;
    dim3 grid((npix + 255) / 256, nviews);
    if (g.kind == MBX_GEOM_TRANSLATION) {
        if (g.rh == 1)
            hipLaunchKernelGGL((pv_table_kernel<MBX_GEOM_TRANSLATION, 1>), grid, dim3(256), 0, st, g, angles, pix, npix,
                               pstride, views, T, NC);
        else
            hipLaunchKernelGGL((pv_table_kernel<MBX_GEOM_TRANSLATION, 2>), grid, dim3(256), 0, st, g, angles, pix, npix,
                               pstride, views, T, NC);
    } else if (g.kind == MBX_GEOM_CONE) {
        if (g.rh == 1)
            hipLaunchKernelGGL((pv_table_kernel<MBX_GEOM_CONE, 1>), grid, dim3(256), 0, st, g, angles, pix, npix,
                               pstride, views, T, NC);
        else
            hipLaunchKernelGGL((pv_table_kernel<MBX_GEOM_CONE, 2>), grid, dim3(256), 0, st, g, angles, pix, npix,
                               pstride, views, T, NC);
    } else {
        if (g.rh == 1)
            hipLaunchKernelGGL((pv_table_kernel<MBX_GEOM_PARALLEL, 1>), grid, dim3(256), 0, st, g, angles, pix, npix,
                               pstride, views, T, NC);
        else
            hipLaunchKernelGGL((pv_table_kernel<MBX_GEOM_PARALLEL, 2>), grid, dim3(256), 0, st, g, angles, pix, npix,
                               pstride, views, T, NC);
    }
    MBX_LAUNCH_CHECK();
    return MBX_OK;
}

template <typename IT, typename OT>
__global__ __launch_bounds__(256) void transpose_kernel(const IT *__restrict__ src, OT *__restrict__ dst, int A,
                                                        int B) {
    // src view: (A, B) row-major -> dst view: (B, A) row-major, one view per blockIdx.z
    __shared__ float t[32][33];
    const size_t vo = (size_t)blockIdx.z * A * B;
    const int a0 = blockIdx.y * 32, b0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int a = a0 + r, bb = b0 + tx;
        if (a < A && bb < B) t[r][tx] = to_float(src[vo + (size_t)a * B + bb]);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int bb = b0 + r, a = a0 + tx;
        if (a < A && bb < B) dst[vo + (size_t)bb * A + a] = from_float<OT>(t[tx][r]);
    }
}

template <typename IT, typename OT>
__global__ __launch_bounds__(256) void convert_kernel(const IT *__restrict__ src, OT *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = from_float<OT>(to_float(src[i]));
}

template <typename IT, typename OT>
static void launch_transpose_t(const void *src, void *dst, int V, int A, int B, bool same_layout, hipStream_t st) {
    if (same_layout) {
        const int64_t n = (int64_t)V * A * B;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL((convert_kernel<IT, OT>), dim3((unsigned)blocks), dim3(256), 0, st, (const IT *)src,
                           (OT *)dst, n);
    } else {
        dim3 grid((B + 31) / 32, (A + 31) / 32, V);
        hipLaunchKernelGGL((transpose_kernel<IT, OT>), grid, dim3(256), 0, st, (const IT *)src, (OT *)dst, A, B);
    }
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int mbx_version(void) { return 1; }

const char *mbx_last_error(void) { return g_last_error.c_str(); }

int64_t mbx_back_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t nviews) {
    DevGeom d;
    if (make_dev_geom(g, &d) != MBX_OK || npix < 0 || nviews < 0) return -1;
    return back_workspace_bytes(d, npix, nviews);
}

int64_t mbx_forward_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t nviews) {
    DevGeom d;
    if (make_dev_geom(g, &d) != MBX_OK || npix < 0 || nviews < 0) return -1;
    return fwd_workspace_bytes(d, npix, nviews);
}

int mbx_forward_project(const mbx_geom *g, const float *angles, const float *vox, const int32_t *pix, int32_t npix,
                        const int32_t *views, int32_t nviews, void *sino, int32_t sino_dtype, int32_t sino_layout,
                        int32_t accumulate, const float *weights, double *sumsq, void *workspace,
                        int64_t workspace_bytes, void *stream) {
    DevGeom d;
    int rc = make_dev_geom(g, &d);
    if (rc != MBX_OK) return rc;
    if (npix < 0) return fail(MBX_EINVAL, "npix must be >= 0");
    if (!views) nviews = d.V;
    if (nviews < 0) return fail(MBX_EINVAL, "nviews must be >= 0");
    if (!dtype_fwd_ok(sino_dtype)) return fail(MBX_EINVAL, "unknown sinogram dtype");
    if (sino_layout != MBX_LAYOUT_VCR && sino_layout != MBX_LAYOUT_VRC)
        return fail(MBX_EINVAL, "unknown sinogram layout");
    if (!angles || !sino || (npix > 0 && (!vox || !pix)))
        return fail(MBX_EINVAL, "NULL device pointer (angles, sino, vox or pix)");
    const int64_t need = fwd_workspace_bytes(d, npix, nviews);
    if (workspace_bytes < need || (need > 0 && !workspace))
        return fail(MBX_EWORKSPACE, "forward workspace too small: need " + std::to_string(need) + " bytes");
    return launch_fwd(d, angles, vox, pix, npix, views, nviews, sino, sino_dtype, sino_layout, accumulate, weights,
                      sumsq, workspace, (hipStream_t)stream);
}

int mbx_forward_plan(const mbx_geom *g, const float *angles, const int32_t *pix, int32_t npix, const int32_t *views,
                     int32_t nviews, void *plan, int64_t plan_bytes, void *stream) {
    DevGeom d;
    int rc = make_dev_geom(g, &d);
    if (rc != MBX_OK) return rc;
    if (npix < 0) return fail(MBX_EINVAL, "npix must be >= 0");
    if (!views) nviews = d.V;
    if (nviews < 0) return fail(MBX_EINVAL, "nviews must be >= 0");
    const int chunk = table_chunk(d, std::max(npix, 1), std::max(nviews, 1));
    if (npix > chunk)
        return fail(MBX_EUNSUPPORTED, "forward plan: " + std::to_string(npix) + " pixels exceed one table chunk (" +
                                          std::to_string(chunk) + "); use mbx_forward_project");
    if (!angles || (npix > 0 && !pix)) return fail(MBX_EINVAL, "NULL device pointer (angles or pix)");
    const int64_t need = fwd_workspace_bytes(d, npix, nviews);
    if (plan_bytes < need || (need > 0 && !plan))
        return fail(MBX_EWORKSPACE, "forward plan too small: need " + std::to_string(need) + " bytes");
    if (nviews == 0) return MBX_OK;
    hipStream_t st = (hipStream_t)stream;
    float *T = (float *)plan;
    int *NC = table_nc(T, chunk, nviews);
    void *aux = (char *)plan + table_bytes(chunk, nviews);
    if (npix > 0) {
        rc = launch_table(d, angles, pix, npix, views, nviews, T, NC, st);
        if (rc != MBX_OK) return rc;
    }
    return launch_fwd_tab(d, T, NC, nullptr, npix, nviews, nullptr, MBX_DTYPE_F32, MBX_LAYOUT_VCR, 0, nullptr, aux,
                          chunk, nullptr, st, 1, kFwdBin);
}

int mbx_forward_apply(const mbx_geom *g, const void *plan, int64_t plan_bytes, const float *vox, int32_t npix,
                      const int32_t *views, int32_t nviews, void *sino, int32_t sino_dtype, int32_t sino_layout,
                      int32_t accumulate, const float *weights, double *sumsq, void *stream) {
    DevGeom d;
    int rc = make_dev_geom(g, &d);
    if (rc != MBX_OK) return rc;
    if (npix < 0) return fail(MBX_EINVAL, "npix must be >= 0");
    if (!views) nviews = d.V;
    if (nviews < 0) return fail(MBX_EINVAL, "nviews must be >= 0");
    if (!dtype_fwd_ok(sino_dtype)) return fail(MBX_EINVAL, "unknown sinogram dtype");
    if (sino_layout != MBX_LAYOUT_VCR && sino_layout != MBX_LAYOUT_VRC)
        return fail(MBX_EINVAL, "unknown sinogram layout");
    const int chunk = table_chunk(d, std::max(npix, 1), std::max(nviews, 1));
    if (npix > chunk) return fail(MBX_EUNSUPPORTED, "forward apply: the pixel set exceeds one table chunk");
    if (!plan || !sino || (npix > 0 && !vox)) return fail(MBX_EINVAL, "NULL device pointer (plan, sino or vox)");
    // the plan's layout follows from its pixel and view counts, and its size
    // grows with either: an apply of more pixels or views than the plan was
    // built for, which would read past its end, is rejected here (counts that
    // differ only within the 256-byte rounding of a tiny plan are not told apart)
    const int64_t need = fwd_workspace_bytes(d, npix, nviews);
    if (plan_bytes != need)
        return fail(MBX_EINVAL, "forward apply: plan_bytes " + std::to_string(plan_bytes) + " is not the size " +
                                    std::to_string(need) + " of a plan of " + std::to_string(npix) + " pixels and " +
                                    std::to_string(nviews) + " views");
    hipStream_t st = (hipStream_t)stream;
    if (nviews == 0) {
        if (sumsq) return launch_sum_partials(nullptr, 0, sumsq, 0, st);  // writes 0
        return MBX_OK;
    }
    const float *T = (const float *)plan;
    const int *NC = table_nc((float *)plan, chunk, nviews);
    void *aux = (char *)plan + table_bytes(chunk, nviews);
    return launch_fwd_tab(d, T, NC, vox, npix, nviews, sino, sino_dtype, sino_layout, accumulate, weights, aux, chunk,
                          sumsq, st, 1, kFwdProject);
}

int mbx_back_project(const mbx_geom *g, const float *angles, const void *sino, int32_t sino_dtype,
                     int32_t num_sino_views, const float *weights, const int32_t *pix, int32_t npix,
                     const int32_t *views, int32_t nviews, int32_t coeff_power, float scale, int32_t accumulate,
                     float *out, void *workspace, int64_t workspace_bytes, void *stream) {
    DevGeom d;
    int rc = make_dev_geom(g, &d);
    if (rc != MBX_OK) return rc;
    if (npix < 0) return fail(MBX_EINVAL, "npix must be >= 0");
    if (!dtype_ok(sino_dtype)) return fail(MBX_EINVAL, "unknown sinogram dtype");
    if (coeff_power != 1 && coeff_power != 2) return fail(MBX_EINVAL, "coeff_power must be 1 or 2");
    if (!views) {
        if (num_sino_views != d.V)
            return fail(MBX_EINVAL, "without view indices the sinogram must hold all " + std::to_string(d.V) +
                                        " views (got " + std::to_string(num_sino_views) + ")");
        nviews = d.V;
    }
    if (nviews < 0) return fail(MBX_EINVAL, "nviews must be >= 0");
    if (npix > 0 && (!angles || !sino || !pix || !out))
        return fail(MBX_EINVAL, "NULL device pointer (angles, sino, pix or out)");
    const int64_t need = back_workspace_bytes(d, npix, nviews);
    if (workspace_bytes < need || (need > 0 && !workspace))
        return fail(MBX_EWORKSPACE, "back workspace too small: need " + std::to_string(need) + " bytes");
    if (nviews == 0) {
        if (!accumulate && npix > 0)
            MBX_HIP_CHECK(hipMemsetAsync(out, 0, (size_t)npix * d.Nz * sizeof(float), (hipStream_t)stream));
        return MBX_OK;
    }
    return launch_back(d, angles, sino, sino_dtype, weights, pix, npix, views, nviews, coeff_power, scale, accumulate,
                       out, workspace, (hipStream_t)stream);
}

int64_t mbx_back_apply_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t nviews) {
    DevGeom d;
    if (make_dev_geom(g, &d) != MBX_OK || npix < 0 || nviews < 0) return -1;
    if (npix == 0 || nviews == 0) return 0;
    return back_part_bytes(d, table_chunk(d, npix, nviews), nviews);
}

// mbx_back_project of a planned pixel set: the geometry table is the plan's.
int mbx_back_apply(const mbx_geom *g, const void *plan, int64_t plan_bytes, const void *sino, int32_t sino_dtype,
                   int32_t num_sino_views, const float *weights, int32_t npix, const int32_t *views, int32_t nviews,
                   int32_t coeff_power, float scale, int32_t accumulate, float *out, void *workspace,
                   int64_t workspace_bytes, void *stream) {
    DevGeom d;
    int rc = make_dev_geom(g, &d);
    if (rc != MBX_OK) return rc;
    if (npix < 0) return fail(MBX_EINVAL, "npix must be >= 0");
    if (!dtype_ok(sino_dtype)) return fail(MBX_EINVAL, "unknown sinogram dtype");
    if (coeff_power != 1 && coeff_power != 2) return fail(MBX_EINVAL, "coeff_power must be 1 or 2");
    if (!views) {
        if (num_sino_views != d.V)
            return fail(MBX_EINVAL, "without view indices the sinogram must hold all " + std::to_string(d.V) +
                                        " views (got " + std::to_string(num_sino_views) + ")");
        nviews = d.V;
    }
    if (nviews < 0) return fail(MBX_EINVAL, "nviews must be >= 0");
    const int chunk = table_chunk(d, std::max(npix, 1), std::max(nviews, 1));
    if (npix > chunk) return fail(MBX_EUNSUPPORTED, "back apply: the pixel set exceeds one table chunk");
    if (npix > 0 && (!plan || !sino || !out)) return fail(MBX_EINVAL, "NULL device pointer (plan, sino or out)");
    // as in mbx_forward_apply: the plan must be the one of npix pixels and nviews views
    const int64_t plan_need = fwd_workspace_bytes(d, npix, nviews);
    if (plan_bytes != plan_need)
        return fail(MBX_EINVAL, "back apply: plan_bytes " + std::to_string(plan_bytes) + " is not the size " +
                                    std::to_string(plan_need) + " of a plan of " + std::to_string(npix) +
                                    " pixels and " + std::to_string(nviews) + " views");
    const int64_t need = mbx_back_apply_workspace_bytes(g, npix, nviews);
    if (workspace_bytes < need || (need > 0 && !workspace))
        return fail(MBX_EWORKSPACE, "back apply workspace too small: need " + std::to_string(need) + " bytes");
    hipStream_t st = (hipStream_t)stream;
    if (npix == 0) return MBX_OK;
    if (nviews == 0) {
        if (!accumulate) MBX_HIP_CHECK(hipMemsetAsync(out, 0, (size_t)npix * d.Nz * sizeof(float), st));
        return MBX_OK;
    }
    return launch_back_tab(d, (const float *)plan, npix, sino, sino_dtype, weights, views, nviews, coeff_power, scale,
                           accumulate, out, (float *)workspace, st, nullptr, table_chunk(d, npix, nviews));
}

int mbx_sino_transpose(const void *src, int32_t src_dtype, int32_t src_layout, void *dst, int32_t dst_dtype,
                       int32_t dst_layout, int32_t num_views, int32_t num_rows, int32_t num_channels, void *stream) {
    if (!dtype_ok(src_dtype) || !dtype_ok(dst_dtype)) return fail(MBX_EINVAL, "unknown dtype");
    if ((src_layout != MBX_LAYOUT_VCR && src_layout != MBX_LAYOUT_VRC) ||
        (dst_layout != MBX_LAYOUT_VCR && dst_layout != MBX_LAYOUT_VRC))
        return fail(MBX_EINVAL, "unknown layout");
    if (num_views < 0 || num_rows <= 0 || num_channels <= 0) return fail(MBX_EINVAL, "bad sinogram shape");
    if (!src || !dst) return fail(MBX_EINVAL, "NULL pointer");
    if (src == dst) return fail(MBX_EINVAL, "src and dst must not alias");
    if (num_views == 0) return MBX_OK;
    // source viewed as (A, B) per view
    const int A = src_layout == MBX_LAYOUT_VRC ? num_rows : num_channels;
    const int B = src_layout == MBX_LAYOUT_VRC ? num_channels : num_rows;
    const bool same = src_layout == dst_layout;
    hipStream_t st = (hipStream_t)stream;
    const bool f16_src = src_dtype == MBX_DTYPE_F16, f16_dst = dst_dtype == MBX_DTYPE_F16;
    if ((f16_src && dst_dtype == MBX_DTYPE_BF16) || (f16_dst && src_dtype == MBX_DTYPE_BF16))
        return fail(MBX_EUNSUPPORTED, "no direct float16 <-> bfloat16 conversion; go through float32");
    if (f16_src && f16_dst)
        launch_transpose_t<f16, f16>(src, dst, num_views, A, B, same, st);
    else if (f16_src)
        launch_transpose_t<f16, float>(src, dst, num_views, A, B, same, st);
    else if (f16_dst)
        launch_transpose_t<float, f16>(src, dst, num_views, A, B, same, st);
    else if (src_dtype == MBX_DTYPE_F32 && dst_dtype == MBX_DTYPE_F32)
        launch_transpose_t<float, float>(src, dst, num_views, A, B, same, st);
    else if (src_dtype == MBX_DTYPE_F32)
        launch_transpose_t<float, bf16>(src, dst, num_views, A, B, same, st);
    else if (dst_dtype == MBX_DTYPE_F32)
        launch_transpose_t<bf16, float>(src, dst, num_views, A, B, same, st);
    else
        launch_transpose_t<bf16, bf16>(src, dst, num_views, A, B, same, st);
    MBX_LAUNCH_CHECK();
    return MBX_OK;
}

}  // extern "C"
