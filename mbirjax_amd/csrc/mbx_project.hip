// Sparse forward / back projectors for the parallel-beam and cone-beam
// geometries on gfx950, plus the sinogram layout transpose.
//
// Operator (identical in both directions, so back == forward^T exactly):
//   A[v, m, n ; p, k] = Ac(p,v,n) * Lr(t_{p,v,k} - m) / cos(phi_{p,v,k})
// with the horizontal fan weight Ac (mbirjax/cone_beam.py:244-284, 409-451,
// 593-644), the vertical fan weight Lr over the projected slice position
// t_k = a + W*k (mbirjax/cone_beam.py:286-377, 480-591) and, for parallel beam,
// Lr = [m == k], cos(phi) = 1 (mbirjax/parallel_beam.py:138-263).
//
// Back projection (Projectors.sparse_back_project, mbirjax/projectors.py:131-209)
//   one wavefront per (pixel, slice chunk, view group); for every view the
//   wave gathers the pixel's 2r+1 channel columns (contiguous rows in the
//   channel-major sinogram) into an LDS detector column, then every lane
//   resamples it onto its slices.  View groups are mapped to XCDs
//   (blockIdx % 8) so each XCD's L2 streams only its share of the sinogram;
//   the group partials are reduced in a fixed order (deterministic).
// Forward projection (Projectors.sparse_forward_project, projectors.py:61-129)
//   one workgroup per (view, channel block, row block) owns an LDS tile of the
//   output; it compacts the subset pixels whose footprint meets the tile (in
//   pixel order), and each wave owns 64 detector rows of the tile, so tile
//   accumulation is race-free and deterministic.  The tile is written once.

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
//   the tap radius and outside [0, C)
//   vertical fan, relative to the reference slice kref = Nz/2 and the integer
//   row ai = rint(t_kref):  t_k - ai = acr + (W + W_lo) * (k - kref)
//   [6] acr   [7] W (rounded)   [8] W_lo = W - float(W)   [9] (W+1)/2
//   [10] min(1, W)   [11] 1/W   [12] u at kref (v/sdd)   [13] du/dk
//   [14] ai (int bits)   [15] first | last row touched by any slice << 16
// plus n_c again in a dense (view, pixel) int array for the forward scan.
// Keeping the row positions relative to ai lets the kernels form t_k - m
// without rounding at the magnitude of a detector row index.

constexpr int kRec = 16;
constexpr int kFill = 4;  // 64-element groups loaded per round (forward staging)
constexpr int kBackFill = 3;  // 64-row chunks loaded per round (back projection)
constexpr int64_t kTableBudget = 512ll << 20;

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
    const int p = pix[(size_t)s * pstride];
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
    const int nc = (int)rint(np);
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

// Rows [first, last] packed in record word 15.
__device__ __forceinline__ int rec_mlo(float w) { return __float_as_int(w) & 0xffff; }
__device__ __forceinline__ int rec_mhi(float w) { return __float_as_int(w) >> 16; }

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

// ======================================================================
// Back projection
// ======================================================================

BackPlan plan_back(const DevGeom &g, int npix, int nviews) {
    BackPlan p;
    const int per_lane = (g.Nz + kWave - 1) / kWave;
    p.kpl = per_lane <= 2 ? 2 : per_lane <= 4 ? 4 : per_lane <= 8 ? 8 : 16;
    p.nchunk = (g.Nz + kWave * p.kpl - 1) / (kWave * p.kpl);
    const int64_t units = ((int64_t)npix + 1) / 2 * p.nchunk;  // waves per view group (two pixels each)
    // Small subsets: split the views over the XCDs so that each XCD's L2
    // streams only its share of the sinogram, keeping every wave of the pass
    // resident at once (7 waves per SIMD); large passes are VALU-bound.  The
    // split (a multiple of the 8 XCDs) grows while the pass has fewer than
    // 4096 waves, every group keeps 4 views or more and the partials stay
    // within 64 MiB.
    if (nviews < 8 || units > 16384) {
        p.nsplit = 1;
    } else {
        p.nsplit = 8;
        while (units * p.nsplit < 4096 && p.nsplit < 64 && nviews >= 8 * p.nsplit &&
               2 * p.nsplit * (int64_t)npix * g.Nz * (int64_t)sizeof(float) <= (64ll << 20))
            p.nsplit *= 2;
    }
    // LDS detector-column capacity per pixel: the rows one slice chunk can
    // reach, from the largest pixel magnification in the recon disc (plus
    // taps and margins), padded to whole 64-row chunks; at least the detector
    // height + 2 plus a chunk (the clamped fallback layout).
    const int round = kWave;
    p.cap = g.R + 8 + round;  // the clamped layout: R + 2 rows, up to 3 more for 4-row alignment, a chunk
    p.unit = 0;
    if (g.kind == MBX_GEOM_CONE) {  // (translation: the clamped layout covers every magnification)
        const double rmax = g.dv * std::sqrt((double)g.img_row_center * g.img_row_center +
                                             (double)g.img_col_center * g.img_col_center) + g.dv;
        const double den = (double)g.inv_M - rmax * (double)g.inv_sdd;
        if (den > 0) {
            const double wmax = g.dv_over_dr / den;
            p.unit = g.rh == 1 && wmax <= 2.0;
            const double want = std::ceil((kWave * p.kpl - 1) * wmax) + 2.0 * g.rh + 8.0 + round;
            if (want < 16384.0) p.cap = std::max(p.cap, (int)want);
        }
    }
    p.cap = (p.cap + 63) & ~63;
    return p;
}

// View-group partials of the plan of `chunk` pixels.
int64_t back_part_bytes(const DevGeom &g, int chunk, int nviews) {
    const BackPlan p = plan_back(g, chunk, nviews);
    const int64_t part = p.nsplit > 1 ? (int64_t)p.nsplit * chunk * g.Nz * (int64_t)sizeof(float) : 0;
    return (part + 255) & ~255ll;
}

// The largest back_part_bytes() of any pixel count 1..n.  A plan's view split
// never grows with the pixel count (more waves stop the doubling earlier), so
// on each run of equal splits the partials grow with the count and the
// maximum sits at the run's last count: found by bisection for each split.
int64_t back_part_bound(const DevGeom &g, int n, int nviews) {
    int64_t best = 0;
    for (int s = 2; s <= 64; s *= 2) {
        if (n < 1 || plan_back(g, 1, nviews).nsplit < s) break;
        int lo = 1, hi = n;  // largest count whose split is >= s
        while (lo < hi) {
            const int mid = lo + (hi - lo + 1) / 2;
            if (plan_back(g, mid, nviews).nsplit >= s) lo = mid; else hi = mid - 1;
        }
        const BackPlan p = plan_back(g, lo, nviews);
        best = std::max<int64_t>(best, (int64_t)p.nsplit * lo * g.Nz * (int64_t)sizeof(float));
    }
    return (best + 255) & ~255ll;
}

int64_t back_workspace_bytes(const DevGeom &g, int npix, int nviews) {
    if (npix == 0 || nviews == 0) return 0;
    const int chunk = table_chunk(g, npix, nviews);
    return table_bytes(chunk, nviews) + back_part_bytes(g, chunk, nviews);
}

// One wavefront per (pixel pair, slice chunk, view group).  For every view the
// wave reads the two pixels' records with scalar loads, gathers each pixel's
// 2r+1 channel columns row-by-row into its LDS detector column (clamped,
// always-valid addresses; rows off the detector are zeroed by a select), then
// every lane resamples both columns onto its slices.  Two pixels per wave keep
// all waves of a subset resident at once, so the waves of an XCD sweep the
// views together and each view is fetched from HBM once.
// The 2r+1 channel columns of one pixel at one view: a buffer resource on the
// view (and on its weights), the byte offset of each column (channels off the
// detector are clamped to a valid column and get weight 0).
template <int RAD>
struct ColumnSource {
    i32x4_t rs, rw;
    int soff[2 * RAD + 1];
    bool cv[2 * RAD + 1];
};

// Vertical fan of one pixel at one view (mbirjax/cone_beam.py:480-541): every
// lane resamples the LDS detector column onto slices k0 + lane + 64 i.
// FIT: the column holds every row the chunk's taps reach (plus a zero or
// don't-care margin row each side), so taps index it directly; otherwise the
// indices are clamped to the stored rows.
// UNIT (coeff_power 1, radius 1, every W <= 2): the clip weights of rows
// mc-1, mc, mc+1 sum to W (a box of width W convolved with a box of width 1,
// sampled at unit spacing), and the side weights (W-1)/2 -+ d never exceed
// W/2 <= min(1, W), so the centre row carries W minus the two sides.
template <int KPL, int RAD, int POW, bool TAIL, bool FIT, bool UNIT>
__device__ __forceinline__ void back_vertical(const float *col, int shift, int n_rows, double acd, double Wd,
                                              float Wf, float hv, float Lv, float c0, float c1, float kr0,
                                              float krlast, float (&acc)[KPL]) {
    // Slice k sits at row ai + t with t = acr + W*(k - kref), formed in FP64 so
    // that the tap distances d = t - rint(t) carry no float32 rounding at the
    // magnitude of a row index; rint is the 1.5*2^52 shift, whose low word is
    // the rounded integer.
    constexpr double kMagic = 6755399441055744.0;
    const float *colp = col - shift;  // colp[m - RAD] is row ai + m (32-bit LDS address arithmetic)
    const float hm1 = hv - 1.0f;
    const double W64 = Wd * (double)kWave;
    const float u0 = fmaf(c1, kr0, c0), c164 = c1 * (float)kWave;
    double t = fma(Wd, (double)kr0, acd);
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        float u;
        if (TAIL) {  // lanes past the last slice repeat it (their sums are never stored)
            const float kr = fminf(kr0 + (float)(kWave * i), krlast);
            t = fma(Wd, (double)kr, acd);
            u = fmaf(c1, kr, c0);
        } else {
            if (i > 0) t += W64;  // slices 64 apart: one FP64 add
            u = fmaf(c164, (float)i, u0);
        }
        const double sh = t + kMagic;
        const int mi = (int)(uint32_t)__builtin_bit_cast(uint64_t, sh);
        const float d = (float)(t - (sh - kMagic));  // |d| <= 1/2
        float sum;
        if (RAD == 1) {
            float cm, cc, cp;
            if (FIT) {
                const float *q = colp + mi;  // one address, three immediate offsets
                cm = q[0];
                cc = q[1];
                cp = q[2];
            } else {
                const int base = mi - shift;
                cm = col[min(max(base, 0), n_rows - 1)];
                cc = col[min(max(base + 1, 0), n_rows - 1)];
                cp = col[min(max(base + 2, 0), n_rows - 1)];
            }
            if (UNIT) {
                const float wm = __builtin_amdgcn_fmed3f(hm1 - d, 0.0f, 1.0f);  // row mc-1
                const float wp = __builtin_amdgcn_fmed3f(hm1 + d, 0.0f, 1.0f);  // row mc+1
                sum = fmaf(wm, cm - cc, fmaf(wp, cp - cc, Wf * cc));
            } else {
                float wm = __builtin_amdgcn_fmed3f(hm1 - d, 0.0f, Lv);        // row mc-1
                float w0 = __builtin_amdgcn_fmed3f(hv - fabsf(d), 0.0f, Lv);  // row mc
                float wp = __builtin_amdgcn_fmed3f(hm1 + d, 0.0f, Lv);        // row mc+1
                if (POW == 2) {
                    w0 *= w0;
                    wp *= wp;
                    wm *= wm;
                }
                sum = w0 * cc;
                sum = fmaf(wp, cp, sum);
                sum = fmaf(wm, cm, sum);
            }
        } else {
            sum = 0.0f;
            const float *q = colp + mi;
#pragma unroll
            for (int j = 0; j <= 2 * RAD; ++j) {
                float wj = __builtin_amdgcn_fmed3f(hv - fabsf(d - (float)(j - RAD)), 0.0f, Lv);
                if (POW == 2) wj *= wj;
                const float cj = FIT ? q[j] : col[min(max(mi - shift + j, 0), n_rows - 1)];
                sum = fmaf(wj, cj, sum);
            }
        }
        const float s2 = fmaf(u, u, 1.0f);
        const float f = (POW == 2) ? s2 : __builtin_amdgcn_sqrtf(s2);
        acc[i] = fmaf(sum, f, acc[i]);
    }
}

// back_kernel flags
constexpr int kBackUnit = 1;  // the unit-sum row weights apply (every W <= 2, radius-1 row taps)
constexpr int kBackVec4 = 2;  // R % 4 == 0, sinogram aligned to 4 elements: 4-row column loads

template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT, bool TAIL>
__device__ __forceinline__ void back_pixel_view(const DevGeom &g, const float *__restrict__ rec,
                                                const ET *__restrict__ sino, const float *__restrict__ w, int vd,
                                                int k0, int kend, int cap, int flags, float *col, int lane,
                                                float (&acc)[KPL]) {
    const int nc = __float_as_int(rec[0]);
    float Aw[2 * RAD + 1];
#pragma unroll
    for (int j = 0; j <= 2 * RAD; ++j) {
        const float a = rec[3 - RAD + j];
        Aw[j] = (POW == 2) ? a * a : a;
    }
    ColumnSource<RAD> src;
    const size_t vofs = (size_t)vd * g.C * g.R;
    src.rs = make_rsrc(sino + vofs);
    if (WGT) src.rw = make_rsrc(w + vofs);
#pragma unroll
    for (int j = 0; j <= 2 * RAD; ++j) {
        const int n = nc - RAD + j;
        src.cv[j] = (n >= 0) && (n < g.C);
        src.soff[j] = (src.cv[j] ? n : 0) * g.R * (int)sizeof(ET);
    }
    if (GEOM == MBX_GEOM_CONE) {
        const float hv = rec[9], Lv = rec[10], c0 = rec[12], c1 = rec[13];
        const int ai = __float_as_int(rec[14]);
        const double Wd = (double)rec[7] + (double)rec[8], acd = (double)rec[6];
        const int mfirst = ai + (int)rint(fma(Wd, (double)(k0 - g.kref), acd)) - RAD;
        const int mlast = ai + (int)rint(fma(Wd, (double)(kend - 1 - g.kref), acd)) + RAD;
        const int mlo = max(mfirst, 0), mhi = min(mlast, g.R - 1);
        if (mlo > mhi) return;  // wave-uniform: this view misses the detector
        // One zero row of margin on each side: a lane's slice position may
        // round to one row beyond the chunk's first or last centre row.
        const int count = mlast - mfirst + 3;
        const bool fit = count + kWave + 3 <= cap;
        int mstart = fit ? mfirst - 1 : mlo - 1;
        int n_rows = fit ? count : mhi - mlo + 3;
        // 4-row column loads (16 bytes float32, 8 bytes bfloat16; R % 4 == 0,
        // aligned sinogram): the column starts on a multiple of 4 rows, so no
        // 4-row group straddles row 0 or row R
        constexpr bool VEC_OK = !WGT;
        const bool v4 = VEC_OK && (flags & kBackVec4);
        if (v4) {
            n_rows += mstart & 3;
            mstart &= ~3;
        }
        // ---- detector column col[i] = row mstart + i: whole 256-row groups
        // 4 rows per lane (v4), the rest in 64-row chunks (loads of up to
        // kFill chunks in flight).  Rows are clamped to the detector for the
        // loads; rows off it are zeroed afterwards.
        const int ng = v4 ? n_rows >> 8 : 0;
        for (int gq = 0; gq < ng; ++gq) {
            const int r = min(max(mstart + gq * 4 * kWave + 4 * lane, 0), g.R - 4);
            mbx_f32x4 ev[2 * RAD + 1];
#pragma unroll
            for (int j = 0; j <= 2 * RAD; ++j) ev[j] = buf_load4<ET>(src.rs, r * (int)sizeof(ET), src.soff[j]);
            mbx_f32x4 c;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float cq = 0.0f;
#pragma unroll
                for (int j = 0; j <= 2 * RAD; ++j) cq = fmaf(src.cv[j] ? Aw[j] : 0.0f, ev[j][q], cq);
                c[q] = cq;
            }
            *(mbx_f32x4 *)(col + gq * 4 * kWave + 4 * lane) = c;  // 16-byte aligned
        }
        constexpr int FILL = kBackFill;
        const int nch = (n_rows + kWave - 1) / kWave;
        for (int c0 = 4 * ng; c0 < nch; c0 += FILL) {
            float ev[FILL][2 * RAD + 1];
#pragma unroll
            for (int u = 0; u < FILL; ++u) {
                if (c0 + u < nch) {  // wave-uniform
                    const int r = min(max(mstart + (c0 + u) * kWave + lane, 0), g.R - 1);
                    const int vo = r * (int)sizeof(ET);
#pragma unroll
                    for (int j = 0; j <= 2 * RAD; ++j) {
                        float e = buf_load<ET>(src.rs, vo, src.soff[j]);
                        if (WGT) e *= buf_load<float>(src.rw, r * 4, src.soff[j] * (int)(4 / sizeof(ET)));
                        ev[u][j] = e;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < FILL; ++u) {
                if (c0 + u < nch) {
                    float c = 0.0f;
#pragma unroll
                    for (int j = 0; j <= 2 * RAD; ++j) c = fmaf(src.cv[j] ? Aw[j] : 0.0f, ev[u][j], c);
                    col[(c0 + u) * kWave + lane] = c;  // cap holds whole chunks
                }
            }
        }
        if (mstart < 0)  // rows above the detector
            for (int i = lane; i < -mstart; i += kWave) col[i] = 0.0f;
        if (mstart + n_rows > g.R)  // rows below it
            for (int i = g.R - mstart + lane; i < n_rows; i += kWave) col[i] = 0.0f;
        wave_sync();
        const int shift = RAD + mstart - ai;
        const float kr0 = (float)(k0 + lane - g.kref), krl = (float)(kend - 1 - g.kref);
        const float Wf = rec[7];
        constexpr bool CAN_UNIT = RAD == 1 && POW == 1;
        if (CAN_UNIT && (flags & kBackUnit)) {  // wave-uniform
            if (fit)
                back_vertical<KPL, RAD, POW, TAIL, true, CAN_UNIT>(col, shift, n_rows, acd, Wd, Wf, hv, Lv, c0, c1,
                                                                   kr0, krl, acc);
            else
                back_vertical<KPL, RAD, POW, TAIL, false, CAN_UNIT>(col, shift, n_rows, acd, Wd, Wf, hv, Lv, c0, c1,
                                                                    kr0, krl, acc);
        } else if (fit) {
            back_vertical<KPL, RAD, POW, TAIL, true, false>(col, shift, n_rows, acd, Wd, Wf, hv, Lv, c0, c1, kr0, krl,
                                                            acc);
        } else {
            back_vertical<KPL, RAD, POW, TAIL, false, false>(col, shift, n_rows, acd, Wd, Wf, hv, Lv, c0, c1, kr0,
                                                             krl, acc);
        }
        wave_sync();
    } else {
#pragma unroll
        for (int i = 0; i < KPL; ++i) {
            const int k = min(k0 + lane + kWave * i, kend - 1);
            float c = 0.0f;
#pragma unroll
            for (int j = 0; j <= 2 * RAD; ++j) {
                float e = buf_load<ET>(src.rs, k * (int)sizeof(ET), src.soff[j]);
                if (WGT) e *= buf_load<float>(src.rw, k * 4, src.soff[j] * (int)(4 / sizeof(ET)));
                c = fmaf(src.cv[j] ? Aw[j] : 0.0f, e, c);
            }
            acc[i] += c;
        }
    }
}

// TAIL: the launch covers only the last, partial slice chunk (lanes past the
// last slice repeat it); otherwise every chunk of the launch is full.
template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT, bool TAIL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 8))) void back_kernel(
    DevGeom g, const float *__restrict__ T, const ET *__restrict__ sino, const float *__restrict__ w, int npix,
    const int32_t *__restrict__ views, int nviews, int nsplit, int chunk0, int nchunk, int cap, int flags, float scale,
    int accumulate, float *__restrict__ out, float *__restrict__ part) {
    constexpr int PPW = 2;  // pixels per wave
    extern __shared__ float lds[];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int split = blockIdx.x % nsplit;
    const int grp = blockIdx.x / nsplit;
    const int chunk = chunk0 + grp % nchunk;
    const int q = __builtin_amdgcn_readfirstlane((grp / nchunk) * 4 + wave);
    const int sA = PPW * q;
    if (sA >= npix) return;  // whole wavefront leaves; no block barriers below
    const bool hasB = sA + 1 < npix;
    const int sB = hasB ? sA + 1 : sA;
    const int k0 = chunk * kWave * KPL;
    const int kend = min(g.Nz, k0 + kWave * KPL);
    const int vb = (int)(((int64_t)nviews * split) / nsplit);
    const int ve = (int)(((int64_t)nviews * (split + 1)) / nsplit);
    float *colA = lds + wave * PPW * cap;
    float *colB = colA + cap;

    float accA[KPL], accB[KPL];
#pragma unroll
    for (int i = 0; i < KPL; ++i) accA[i] = accB[i] = 0.0f;
    for (int vi = vb; vi < ve; ++vi) {
        const int vd = views ? views[vi] : vi;
        const float *recA = T + ((size_t)vi * npix + sA) * kRec;
        const float *recB = T + ((size_t)vi * npix + sB) * kRec;
        back_pixel_view<GEOM, KPL, RAD, ET, POW, WGT, TAIL>(g, recA, sino, w, vd, k0, kend, cap, flags, colA, lane,
                                                            accA);
        if (hasB)
            back_pixel_view<GEOM, KPL, RAD, ET, POW, WGT, TAIL>(g, recB, sino, w, vd, k0, kend, cap, flags, colB, lane,
                                                                accB);
    }
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        const int k = k0 + lane + kWave * i;
        if (k < kend) {
            if (nsplit == 1) {
                const size_t oA = (size_t)sA * g.Nz + k;
                const float vA = scale * accA[i];
                out[oA] = accumulate ? out[oA] + vA : vA;
                if (hasB) {
                    const size_t oB = (size_t)sB * g.Nz + k;
                    const float vB = scale * accB[i];
                    out[oB] = accumulate ? out[oB] + vB : vB;
                }
            } else {
                part[((size_t)split * npix + sA) * g.Nz + k] = accA[i];
                if (hasB) part[((size_t)split * npix + sB) * g.Nz + k] = accB[i];
            }
        }
    }
}

// out = (accumulate ? out : 0) + scale * sum_s part[s]   (fixed order over s)
__global__ __launch_bounds__(256) void back_reduce_kernel(const float *__restrict__ part, int nsplit, int64_t n,
                                                          float scale, int accumulate, float *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int j = 0; j < nsplit; ++j) s += part[(size_t)j * n + i];
        const float v = scale * s;
        out[i] = accumulate ? out[i] + v : v;
    }
}

template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT, bool TAIL>
static void launch_back_range(const DevGeom &g, const BackPlan &pl, const float *T, const ET *sino, const float *w,
                              int npix, const int32_t *views, int nviews, int chunk0, int nch, float scale,
                              int accumulate, float *out, float *part, hipStream_t st) {
    constexpr int PPW = 2;  // pixels per wave, as in back_kernel
    const int64_t groups = ((int64_t)npix + 4 * PPW - 1) / (4 * PPW);  // 4 waves per block
    const int64_t blocks = groups * nch * pl.nsplit;
    const size_t lds = GEOM == MBX_GEOM_CONE ? 4 * PPW * (size_t)pl.cap * sizeof(float) : 0;
    const bool vec4 = g.R % 4 == 0 && ((uintptr_t)sino & (4 * sizeof(ET) - 1)) == 0;
    const int flags = (pl.unit ? kBackUnit : 0) | (vec4 ? kBackVec4 : 0);
    hipLaunchKernelGGL((back_kernel<GEOM, KPL, RAD, ET, POW, WGT, TAIL>), dim3((unsigned)blocks), dim3(256), lds,
                       st, g, T, sino, w, npix, views, nviews, pl.nsplit, chunk0, nch, pl.cap, flags, scale,
                       accumulate, out, part);
}

// Full slice chunks in one launch, a partial last chunk in a second one.
template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT>
static void launch_back_w(const DevGeom &g, const BackPlan &pl, const float *T, const ET *sino, const float *w,
                          int npix, const int32_t *views, int nviews, float scale, int accumulate, float *out,
                          float *part, hipStream_t st) {
    const int full = g.Nz / (kWave * KPL);
    if (full > 0)
        launch_back_range<GEOM, KPL, RAD, ET, POW, WGT, false>(g, pl, T, sino, w, npix, views, nviews, 0, full,
                                                                      scale, accumulate, out, part, st);
    if (full < pl.nchunk)
        launch_back_range<GEOM, KPL, RAD, ET, POW, WGT, true>(g, pl, T, sino, w, npix, views, nviews, full, 1,
                                                                     scale, accumulate, out, part, st);
}

template <int GEOM, int KPL, int RAD, typename ET, int POW>
static void launch_back_t(const DevGeom &g, const BackPlan &pl, const float *T, const ET *sino, const float *w,
                          int npix, const int32_t *views, int nviews, float scale, int accumulate, float *out,
                          float *part, hipStream_t st) {
    if (w)
        launch_back_w<GEOM, KPL, RAD, ET, POW, true>(g, pl, T, sino, w, npix, views, nviews, scale, accumulate,
                                                            out, part, st);
    else
        launch_back_w<GEOM, KPL, RAD, ET, POW, false>(g, pl, T, sino, w, npix, views, nviews, scale, accumulate,
                                                             out, part, st);
}

template <int GEOM, int KPL, int RAD>
static int dispatch_back_3(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, int dt,
                           const float *w, int npix, const int32_t *views, int nviews, int pw, float scale,
                           int accumulate, float *out, float *part, hipStream_t st) {
    if (pw == 2) {
        if (dt != MBX_DTYPE_F32) return fail(MBX_EUNSUPPORTED, "coeff_power=2 supports float32 input only");
        launch_back_t<GEOM, KPL, RAD, float, 2>(g, pl, T, (const float *)sino, w, npix, views, nviews, scale,
                                                accumulate, out, part, st);
    } else if (dt == MBX_DTYPE_F32) {
        launch_back_t<GEOM, KPL, RAD, float, 1>(g, pl, T, (const float *)sino, w, npix, views, nviews, scale,
                                                accumulate, out, part, st);
    } else if (dt == MBX_DTYPE_BF16) {
        launch_back_t<GEOM, KPL, RAD, bf16, 1>(g, pl, T, (const bf16 *)sino, w, npix, views, nviews, scale,
                                               accumulate, out, part, st);
    } else {
        launch_back_t<GEOM, KPL, RAD, f16, 1>(g, pl, T, (const f16 *)sino, w, npix, views, nviews, scale, accumulate,
                                              out, part, st);
    }
    return MBX_OK;
}

template <int GEOM, int KPL>
static int dispatch_back_2(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, int dt,
                           const float *w, int npix, const int32_t *views, int nviews, int pw, float scale,
                           int accumulate, float *out, float *part, hipStream_t st) {
    if (g.rh == 1)
        return dispatch_back_3<GEOM, KPL, 1>(g, pl, T, sino, dt, w, npix, views, nviews, pw, scale, accumulate, out,
                                             part, st);
    return dispatch_back_3<GEOM, KPL, 2>(g, pl, T, sino, dt, w, npix, views, nviews, pw, scale, accumulate, out, part,
                                         st);
}

template <int GEOM>
static int dispatch_back_1(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, int dt,
                           const float *w, int npix, const int32_t *views, int nviews, int pw, float scale,
                           int accumulate, float *out, float *part, hipStream_t st) {
    switch (pl.kpl) {
        case 4:
            return dispatch_back_2<GEOM, 4>(g, pl, T, sino, dt, w, npix, views, nviews, pw, scale, accumulate, out,
                                            part, st);
        case 2:
            return dispatch_back_2<GEOM, 2>(g, pl, T, sino, dt, w, npix, views, nviews, pw, scale, accumulate, out,
                                            part, st);
        case 8:
            return dispatch_back_2<GEOM, 8>(g, pl, T, sino, dt, w, npix, views, nviews, pw, scale, accumulate, out,
                                            part, st);
        default:
            return dispatch_back_2<GEOM, 16>(g, pl, T, sino, dt, w, npix, views, nviews, pw, scale, accumulate, out,
                                             part, st);
    }
}

// Back projection of pixels whose table T (nviews x npix records) is built.
int launch_back_tab(const DevGeom &g, const float *T, int npix, const void *sino, int sino_dtype,
                    const float *weights, const int32_t *views, int nviews, int coeff_power, float scale,
                    int accumulate, float *out, float *part, hipStream_t st, int *defer, int plan_npix) {
    if (defer) *defer = 0;
    if (npix == 0) return MBX_OK;
    const BackPlan pl = plan_back(g, std::max(npix, plan_npix), nviews);
    if (pl.nsplit > 1 && !part) return fail(MBX_EWORKSPACE, "back projection needs a partials workspace");
    int rc;
    if (is_fan(g))
        rc = dispatch_back_1<MBX_GEOM_CONE>(g, pl, T, sino, sino_dtype, weights, npix, views, nviews, coeff_power,
                                            scale, accumulate, out, part, st);
    else
        rc = dispatch_back_1<MBX_GEOM_PARALLEL>(g, pl, T, sino, sino_dtype, weights, npix, views, nviews, coeff_power,
                                                scale, accumulate, out, part, st);
    if (rc != MBX_OK) return rc;
    MBX_LAUNCH_CHECK();
    if (pl.nsplit > 1 && defer) {
        *defer = pl.nsplit;  // the caller sums the partials
        return MBX_OK;
    }
    if (pl.nsplit > 1) {
        const int64_t n = (int64_t)npix * g.Nz;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(back_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, part, pl.nsplit, n, scale,
                           accumulate, out);
        MBX_LAUNCH_CHECK();
    }
    return MBX_OK;
}

int launch_back(const DevGeom &g, const float *angles, const void *sino, int sino_dtype, const float *weights,
                const int32_t *pix, int npix, const int32_t *views, int nviews, int coeff_power, float scale,
                int accumulate, float *out, void *ws, hipStream_t st) {
    if (npix == 0 || nviews == 0) return MBX_OK;
    if (!ws) return fail(MBX_EWORKSPACE, "back projection needs a workspace");
    const int chunk = table_chunk(g, npix, nviews);
    char *w8 = (char *)ws;
    float *T = (float *)w8;
    int *NC = table_nc(T, chunk, nviews);
    float *part = (float *)(w8 + table_bytes(chunk, nviews));
    for (int c0 = 0; c0 < npix; c0 += chunk) {
        const int n = std::min(chunk, npix - c0);
        int rc = launch_table(g, angles, pix + c0, n, views, nviews, T, NC, st);
        if (rc != MBX_OK) return rc;
        // every chunk planned as a full one: the partials were sized for it
        rc = launch_back_tab(g, T, n, sino, sino_dtype, weights, views, nviews, coeff_power, scale, accumulate,
                             out + (size_t)c0 * g.Nz, part, st, nullptr, chunk);
        if (rc != MBX_OK) return rc;
    }
    return MBX_OK;
}

// ======================================================================
// Forward projection
// ======================================================================

constexpr int kFwdCB = 16;        // channels per block, (views, rows, channels) output: the per-lane register tile
constexpr int kFwdCBW = 32;       // channels per block, (views, channels, rows) output: the sliding window's sweep
constexpr int kFwdCBWMin = 8;     // ... narrowed down to this on small launches
constexpr int64_t kFwdMinWaves = 4096;  // four waves per SIMD
constexpr int kFwdRPL = 2;        // detector rows per lane (64 apart), (views, rows, channels) output
constexpr int kFwdRPLW = 4;       // detector rows per lane, (views, channels, rows) output (2 on
                                  // detectors of at most 256 rows)
constexpr int kFwdPad = 8;        // zero slices either side of a staged cylinder (>= 2*bp_psf_radius + 1;
                                  // a multiple of 4 keeps every staged cylinder 16-byte aligned)
constexpr int kFwdMaxBatch = 16;  // cylinders staged per batch
constexpr size_t kFwdLds = 36 * 1024;   // both staging buffers of a block, (views, rows, channels) output
constexpr size_t kFwdLdsW = 20 * 1024;  // the same, (views, channels, rows) output: small batches, more blocks per CU
constexpr int kBinMax = 16384;    // channel buckets the binning kernel holds in LDS

FwdPlan plan_fwd(const DevGeom &g, int nviews, int layout) {
    FwdPlan p;
    p.rpl = layout == MBX_LAYOUT_VCR ? (g.R <= 4 * kWave ? 2 : kFwdRPLW) : kFwdRPL;
    const int slab = kWave * p.rpl;  // rows per wave
    p.nw = (std::min(g.R, 4 * slab) + slab - 1) / slab;
    p.rows_pb = p.nw * slab;
    p.nrb = (g.R + p.rows_pb - 1) / p.rows_pb;
    p.cb = layout == MBX_LAYOUT_VCR ? kFwdCBW : kFwdCB;
    p.ncb = (g.C + p.cb - 1) / p.cb;
    // channel-major output: narrower channel blocks while the launch has
    // fewer than kFwdMinWaves waves, each block then walking fewer candidates
    while (layout == MBX_LAYOUT_VCR && p.cb > kFwdCBWMin &&
           (int64_t)nviews * p.ncb * p.nrb * p.nw < kFwdMinWaves) {
        p.cb /= 2;
        p.ncb = (g.C + p.cb - 1) / p.cb;
    }
    p.list = 0;
    // A staged cylinder holds the slices the block's rows reach (at most all
    // slices) in whole 64-slice chunks, and the zero pads.
    const int span_max = is_fan(g) ? g.Nz : p.rows_pb;
    p.cstride = ((span_max + kWave - 1) / kWave) * kWave + 2 * kFwdPad + 8;  // a multiple of 4
    p.nblocks = (int64_t)nviews * p.ncb * p.nrb;
    size_t budget = layout == MBX_LAYOUT_VCR ? kFwdLdsW : kFwdLds;
    // Launches of few waves (at most two rounds of kFwdMinWaves): a smaller
    // staging batch if that lets every block be resident at once (160 KB of
    // LDS per CU, 1 KB of allocation slack per block), down to 4 slots.
    const size_t slot_bytes = 2 * (size_t)p.cstride * sizeof(float);
    if (p.nblocks * p.nw <= 2 * kFwdMinWaves) {
        const int64_t per_cu = (p.nblocks + 255) / 256;
        const size_t fit = (size_t)(160 * 1024 / per_cu) - 1024 - 64 * sizeof(int);
        if (fit < budget && fit >= 4 * slot_bytes) budget = fit;
    }
    p.nb = (int)std::max<size_t>(1, std::min<size_t>(kFwdMaxBatch, budget / slot_bytes));
    p.lds = 2 * (size_t)p.nb * p.cstride * sizeof(float) + 64 * sizeof(int);
    return p;
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
// FP64 partial slots of the squared norm: one per block of either output layout
static int64_t fwd_part_blocks(const DevGeom &g, int nviews) {
    return std::max(plan_fwd(g, nviews, MBX_LAYOUT_VCR).nblocks, plan_fwd(g, nviews, MBX_LAYOUT_VRC).nblocks);
}
static int fwd_nbk(const DevGeom &g) { return g.C + 2 * g.rh; }  // centre-channel buckets n_c + r
int64_t fwd_blocks(const DevGeom &g, int nviews, int layout) { return plan_fwd(g, nviews, layout).nblocks; }

// Forward record of one (view, pixel) pair, stored in the view's binned order
// (centre channel, then pixel): what the forward update reads per candidate
// pixel, so that a block walks its candidates through contiguous scalar loads.
//   [0..2] Ac(n_c-1 .. n_c+1)   [3] W   [4] W_lo
//   [5] acr (acr - (1-W)/2 in the two-slice mode)   [6] 1/W   [7] -acr/W
//   [8] ai (float)   [9], [10] clamp range of the resampling slice index,
//   relative to kref   [11] kf | span << 16: the slices kf .. kf+span-1 reach
//   a detector row   [12] bit g: a row of 128g .. 128g+127 is reached (all
//   bits when the detector has more than 4096 rows)   [13] (1-W)/2
//   [14] v/sdd at kref   [15] its step per slice
constexpr int kFRec = 16;

// Forward scratch after the geometry table: per-block FP64 partials of the
// squared norm, then per view the pixel list sorted by centre channel, the
// bucket starts and the forward records in that order.
int64_t fwd_aux_bytes(const DevGeom &g, int chunk, int nviews) {
    return (int64_t)(al256((size_t)fwd_part_blocks(g, nviews) * sizeof(double)) +
                     al256((size_t)nviews * chunk * sizeof(int)) +
                     al256((size_t)nviews * (fwd_nbk(g) + 1) * sizeof(int)) +
                     al256((size_t)nviews * chunk * kFRec * sizeof(float)));
}

int64_t fwd_workspace_bytes(const DevGeom &g, int npix, int nviews) {
    if (nviews == 0) return 0;  // an empty pixel set still writes the views (zeros or unchanged)
    const int chunk = table_chunk(g, std::max(npix, 1), nviews);
    return table_bytes(chunk, nviews) + fwd_aux_bytes(g, chunk, nviews);
}

// Vertical resampling of a staged cylinder onto row m (row-driven fan,
// mbirjax/cone_beam.py:347-372).  The weights clip((W+1)/2 - |t_k - m|, 0,
// min(1, W)) of the slices k are a box of width W convolved with a box of
// width 1, sampled at spacing W, so over all k they sum to 1.  Three modes:
//   kVertGeneral: every tap of the bp_psf_radius window with its own weight;
//   kVertUnit:    W >= 1/2 for every pixel (bp_psf_radius 1): the slices
//                 k_r -+ 1 around the nearest k_r carry clip((1-W)/2 -+ dm, 0, 1)
//                 (never above 1/2 <= min(1, W)) and k_r the rest of 1;
//   kVertTwo:     W >= 1 for every pixel: only k0 = floor((m - t_0)/W) and
//                 k0 + 1 are non-zero, the latter with f = clip(m - t_k0 +
//                 (1-W)/2, 0, 1), the former with 1 - f.
// The smallest pixel magnification is bounded through the recon's circumradius.
constexpr int kVertGeneral = 0, kVertUnit = 1, kVertTwo = 2;

static int fwd_vert_mode(const DevGeom &g) {
    if (g.kind != MBX_GEOM_CONE || g.rb != 1) return kVertGeneral;
    const mbx_geom &h = g.h;
    const double rmax = h.delta_voxel * std::sqrt(0.25 * (double)h.num_recon_rows * h.num_recon_rows +
                                                  0.25 * (double)h.num_recon_cols * h.num_recon_cols);
    const double inv_sdd = std::isfinite(h.source_detector_dist) ? 1.0 / h.source_detector_dist : 0.0;
    const double den = 1.0 / h.magnification + rmax * inv_sdd;
    if (!(den > 0)) return kVertGeneral;
    const double wmin = (h.delta_voxel / den) / h.delta_det_row;
    return wmin >= 1.0 ? kVertTwo : wmin >= 0.5 ? kVertUnit : kVertGeneral;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *scratch) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += scratch[i];
    return r;  // valid in thread 0
}

// Per view: the subset pixels sorted by centre channel n_c (bucket n_c + r),
// ascending pixel slot within a bucket, the bucket starts, and the forward
// records in that order.  Pixels whose taps miss the detector are dropped.
// Stable counting sort: wave s (s < nseg) owns the contiguous slot range
// [s*npix/nseg, (s+1)*npix/nseg); pass 1 counts its pixels per bucket, a scan
// over (bucket, segment) gives every segment its cursor in each bucket, and
// pass 2 walks the range 64 slots at a time in order, sorting the round's
// (bucket, lane) keys across the wave and ranking each slot within its bucket,
// so the order -- and the forward summation order -- is ascending slot within
// each bucket whatever the atomics' timing; the position goes back to the
// slot's own lane, which loaded the slot's table record while the round was
// sorted and writes its forward record there.
constexpr int kBinThreads = 1024;
constexpr int kBinWaves = kBinThreads / 64;
constexpr size_t kBinLds = 64 * 1024;  // counters [nseg][nbk]
__global__ __launch_bounds__(kBinThreads) void fwd_bin_kernel(DevGeom g, const float *__restrict__ T,
                                                              const int *__restrict__ NC, int npix, int nbk,
                                                              int vmode, int nseg, int *__restrict__ list,
                                                              int *__restrict__ cstart, float *__restrict__ FR) {
    extern __shared__ int cnt[];  // [nseg][nbk]
    const int rad = g.rh;
    const int v = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int *ncv = NC + (size_t)v * npix;
    int *lg = list + (size_t)v * npix;
    int *cs = cstart + (size_t)v * (nbk + 1);
    for (int i = tid; i < nseg * nbk; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const int sb = (int)((int64_t)npix * wave / nseg), se = (int)((int64_t)npix * (wave + 1) / nseg);
    int *cw = cnt + (size_t)wave * nbk;  // this wave's counters (wave < nseg)
    if (wave < nseg) {  // pass 1
        for (int s = sb + lane; s < se; s += 64) {
            const int b = ncv[s] + rad;
            if (b >= 0 && b < nbk) atomicAdd(&cw[b], 1);
        }
    }
    __syncthreads();
    // exclusive scan over (bucket, segment): each thread owns a contiguous run of buckets
    const int per = (nbk + blockDim.x - 1) / blockDim.x;
    const int b0 = tid * per, b1 = min(nbk, b0 + per);
    int run = 0;
    for (int b = b0; b < b1; ++b)
        for (int q = 0; q < nseg; ++q) run += cnt[q * nbk + b];
    __shared__ int wsum[kBinWaves];
    int incl = run;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int base = incl - run;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    for (int b = b0; b < b1; ++b) {
        cs[b] = base;
        for (int q = 0; q < nseg; ++q) {  // segment q's cursor in bucket b
            const int c = cnt[q * nbk + b];
            cnt[q * nbk + b] = base;
            base += c;
        }
    }
    if (tid == blockDim.x - 1) cs[nbk] = base;
    __syncthreads();
    // forward record of slot s at list position i
    auto write_record = [&](int i, const float4 &R0, const float4 &R1, const float4 &R2, const float4 &R3) {
        float o[kFRec];
        o[0] = R0.z;
        o[1] = R0.w;
        o[2] = R1.x;
        int rlo = 0, rhi = g.R - 1, kf = 0, span = 0;
        constexpr int P = kFwdPad;
        float clo = 0.0f, chi = 0.0f;
        if (is_fan(g)) {
            const float acr = R1.z, W = R1.w, Wlo = R2.x, hv = R2.y, invW = R2.w;
            const int ai = __float_as_int(R3.z);
            rlo = rec_mlo(R3.w);
            rhi = rec_mhi(R3.w);
            // slices whose taps reach rows rlo .. rhi (nearest slice -+ rb)
            kf = max(g.kref + (int)rintf(((float)(rlo - ai) - acr) * invW) - g.rb, 0);
            const int kl = min(g.kref + (int)rintf(((float)(rhi - ai) - acr) * invW) + g.rb, g.Nz - 1);
            span = (rlo <= rhi && kf <= kl) ? kl - kf + 1 : 0;
            if (span == 0) {  // no wave's rows lie in [0xffff, 0]
                kf = 0;
                rlo = 0xffff;
                rhi = 0;
            }
            const float q1 = hv - W;
            o[3] = W;
            o[4] = Wlo;
            o[5] = vmode == kVertTwo ? acr - q1 : acr;
            o[6] = invW;
            o[7] = -(acr * invW);
            o[8] = (float)ai;
            o[13] = q1;
            o[14] = R3.x;
            o[15] = R3.y;
            // staged slice i of the cylinder is slice kf - P + i; slices kf-P ..
            // kf-1 and kf+span .. kf+span+P-1 are zeros.  Clamping the nearest
            // (or lower) slice keeps every tap there; a clamped row only meets zeros.
            if (vmode == kVertTwo) {
                clo = (float)(kf - g.kref - 2);
                chi = (float)(kf - g.kref + span);
            } else {
                clo = (float)(kf - g.kref - P + g.rb);
                chi = (float)(kf - g.kref + span + P - 1 - g.rb);
            }
        } else {
            o[3] = 1.0f;
            o[4] = 0.0f;
            o[5] = 0.0f;
            o[6] = 1.0f;
            o[7] = 0.0f;
            o[8] = 0.0f;
            o[13] = 0.0f;
            o[14] = 0.0f;
            o[15] = 0.0f;
        }
        o[9] = clo;
        o[10] = chi;
        o[11] = __int_as_float(kf | (span << 16));
        unsigned gm = 0;
        if (g.R > 32 * 128) {
            gm = rlo <= rhi ? 0xffffffffu : 0u;
        } else if (rlo <= rhi) {
            for (int q = rlo >> 7; q <= (rhi >> 7); ++q) gm |= 1u << q;
        }
        o[12] = __uint_as_float(gm);
        float4 *dst = (float4 *)(FR + ((size_t)v * npix + i) * kFRec);
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q] = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    };
    if (wave < nseg) {  // pass 2: place in slot order; each lane writes one slot's record
        for (int s0 = sb; s0 < se; s0 += 64) {
            const int s = s0 + lane;
            const int sc = min(s, se - 1);  // (lanes past the range load a valid slot, unused)
            // this lane's slot: its table record is loaded now, while the round is sorted
            const float4 *r4 = (const float4 *)(T + ((size_t)v * npix + sc) * kRec);
            const float4 R0 = r4[0], R1 = r4[1], R2 = r4[2], R3 = r4[3];
            const int b = s < se ? ncv[sc] + rad : -1;
            const bool ok = b >= 0 && b < nbk;
            // sort the round's (bucket, lane) keys across the wave (bitonic
            // network over lane shuffles): equal buckets become adjacent, in lane
            // (= slot) order; slots off the detector sort last
            int key = ((ok ? b : nbk) << 6) | lane;
#pragma unroll
            for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const int other = __shfl_xor(key, j, 64);
                    const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
                    key = keep_min ? min(key, other) : max(key, other);
                }
            }
            const int kb = key >> 6;
            const int ks = s0 + (key & 63);
            // rank within the bucket: lanes since the bucket's first lane (max-scan of heads)
            const int prev = __shfl_up(kb, 1, 64);
            int first = (lane == 0 || prev != kb) ? lane : 0;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int t = __shfl_up(first, d, 64);
                if (lane >= d) first = max(first, t);
            }
            const int next = __shfl_down(kb, 1, 64);
            const bool last = lane == 63 || next != kb;
            int pos = 0;
            if (kb < nbk) {
                const int cur = cw[kb];  // read by every lane of the bucket before its last lane moves it
                pos = cur + lane - first;
                if (last) cw[kb] = pos + 1;
                lg[pos] = ks;
            }
            // back to the slot's own lane (ds_permute: lane j sends to lane key & 63)
            const int mypos = __builtin_amdgcn_ds_permute((key & 63) << 2, pos);
            if (ok) write_record(mypos, R0, R1, R2, R3);
        }
    }
}

// One workgroup per (view, channel block, row block).  The block's candidate
// pixels are a contiguous run of the view's binned records (sorted by centre
// channel, then pixel) and are handled in batches, double-buffered in LDS:
//   staging: the slices of each candidate's cylinder that reach the block's
//            rows are copied global -> LDS asynchronously (buffer loads into
//            LDS) while the previous batch is being used, then scaled in place
//            by 1/cos(phi_k) with zero pads;
//   update:  every wave resamples the staged cylinders onto its rows
//            (row-driven vertical fan, mbirjax/cone_beam.py:347-372) and weights
//            them with the horizontal taps (mbirjax/cone_beam.py:276-282).
// Channel-major output (views, channels, rows): 32-channel blocks (8 or 16 on
// launches of fewer than 4096 waves) of 2 waves,
// lane l of wave w owning rows m0 + 256w + l + 64r (r < 4); the candidates are
// walked in list order and a sliding window of 2*RAD+1 channel accumulators
// per row follows their buckets, streaming each channel out as soon as it is
// complete.  Row-major output (views, rows, channels): 16-channel blocks of 4
// waves, lane l of wave w owning rows m0 + 128w + l and + 64, with the 16
// channels in registers (a compile-time loop over the centre-channel buckets
// puts every tap in a fixed register) and the tile written once.
// Each output element has one writer and a fixed summation order (the list
// order of its candidates, the same in both schemes).
template <int GEOM, int RAD, int RB, int VM, typename OT, int LAYOUT, int RPLW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void fwd_kernel(
    DevGeom g, const float *__restrict__ T, const float *__restrict__ FR, const int *__restrict__ list,
    const int *__restrict__ cstart, const float *__restrict__ vox, int64_t vstride, int npix, int ncb, int nrb,
    int cbw, int nbatch, int cstride, OT *__restrict__ sino, int accumulate, const float *__restrict__ w,
    double *__restrict__ part_sumsq) {
    // (views, channels, rows) output: a sliding window of 2*RAD+1 channels per row
    constexpr bool SW = LAYOUT == MBX_LAYOUT_VCR;
    constexpr int CB = kFwdCB, P = kFwdPad, RPL = SW ? RPLW : kFwdRPL;
    static_assert(RPL % 2 == 0, "rows per lane cover whole 128-row groups of the record masks");
    // channels per block: the register tile's CB, or the window's cbw (at most
    // kFwdCBW, so that the bucket starts fit one per lane)
    const int cbs = SW ? cbw : CB;
    constexpr int NDT = CB + 2 * RAD;  // register tile: centre-channel buckets n0 - RAD .. n0 + CB - 1 + RAD
    const int ND = cbs + 2 * RAD;      // buckets of the block
    static_assert(kFwdCBW + 2 * RAD < kWave, "the bucket starts are held one per lane");
    extern __shared__ float lds[];
    const int nw = blockDim.x >> 6;
    float *bufs = lds;                                                  // [2][nbatch][cstride]
    double *dscratch = (double *)(lds + 2 * (size_t)nbatch * cstride);  // [8]

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int b = blockIdx.x;
    const int vi = b / (ncb * nrb);
    const int rem = b - vi * (ncb * nrb);
    const int cbi = rem / nrb, rbi = rem - (rem / nrb) * nrb;
    const int n0 = cbi * cbs, n1 = min(g.C, n0 + cbs);
    const int rows_pb = nw * kWave * RPL;
    const int m0 = rbi * rows_pb, m1 = min(g.R, m0 + rows_pb);
    const int wm0 = m0 + wave * kWave * RPL;  // this wave's rows: wm0 + lane + 64 r
    const int wml = min(m1, wm0 + kWave * RPL) - 1;
    const float *Tv = T + (size_t)vi * npix * kRec;
    const float *FRv = FR + (size_t)vi * npix * kFRec;
    const int *lv = list + (size_t)vi * npix;  // pixel slot of each record
    const int nbk = g.C + 2 * RAD;
    const int *csv = cstart + (size_t)vi * (nbk + 1);
    // bucket starts of this block: bs[D] (or lane D of bsl) is the start of
    // centre channel n0 + D - RAD
    int bs[SW ? 1 : NDT + 1];
    int bsl = 0, lo, hi;
    if constexpr (SW) {
        bsl = csv[min(n0 + min(lane, ND), nbk)];
        lo = __builtin_amdgcn_readlane(bsl, 0);
        hi = __builtin_amdgcn_readlane(bsl, ND);
    } else {
#pragma unroll
        for (int D = 0; D <= NDT; ++D) bs[D] = csv[min(n0 + D, nbk)];
        lo = bs[0];
        hi = bs[NDT];
    }
    const int nbat = (hi - lo + nbatch - 1) / nbatch;

    // the register tile (CB channels), or the window: acc[r][t] is channel
    // n0 - 2*RAD + dc + t while centre-channel bucket dc is being added
    constexpr int NA = SW ? 2 * RAD + 1 : CB;
    float acc[RPL][NA];
#pragma unroll
    for (int r = 0; r < RPL; ++r)
#pragma unroll
        for (int c = 0; c < NA; ++c) acc[r][c] = 0.0f;

    // Staging data of a batch, lane i holding slot i's: kf | span << 16, pixel
    // slot, v/sdd at kref and its step per slice.  Loaded a batch ahead, without
    // conditions (lanes past the batch load some record of the block and are
    // never read), so that nothing waits for the loads before they are used.
    struct Meta {
        int ks, sc;
        float u0, du;
    };
    auto load_meta = [&](int bi) {
        Meta q;
        const int i = min(lo + bi * nbatch + min(lane, nbatch - 1), hi - 1);
        const float *rf = FRv + (size_t)i * kFRec;
        q.ks = GEOM == MBX_GEOM_CONE ? __float_as_int(rf[11]) : (m0 | ((m1 - m0) << 16));
        q.sc = lv[i];
        q.u0 = rf[14];
        q.du = rf[15];
        return q;
    };
    // asynchronous global -> LDS copy of the raw slices of this wave's slots
    auto issue = [&](const Meta &q, int bi) {
        float *buf = bufs + (size_t)(bi & 1) * nbatch * cstride;
        const int nb = min(nbatch, hi - lo - bi * nbatch);
        for (int slot = wave; slot < nb; slot += nw) {
            const int ks = __builtin_amdgcn_readlane(q.ks, slot);
            const int kf = ks & 0xffff, span = ks >> 16;
            if (span == 0) continue;
            const int sc = __builtin_amdgcn_readlane(q.sc, slot);
            // slices past the cylinder end read 0 (num_records bounds the buffer)
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc((void *)(vox + sc * vstride + kf), 0, (g.Nz - kf) * 4,
                                                  0x00020000);
            float *dst = buf + (size_t)slot * cstride + P;  // 16-byte aligned
            // whole 256-slice blocks 16 bytes per lane, the rest 64 slices at a time
            const int n4 = span >> 8;
            for (int c = 0; c < n4; ++c)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (__attribute__((address_space(3))) void *)(dst + c * 4 * kWave), 16, (c * kWave + lane) * 16, 0,
                    0, 0);
            const int nch = (span + kWave - 1) / kWave;
            for (int c = 4 * n4; c < nch; ++c)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (__attribute__((address_space(3))) void *)(dst + c * kWave), 4, (c * kWave + lane) * 4, 0, 0,
                    0);
        }
    };
    // in-place scaling by 1/cos(phi_k), zero past the span, zero pads
    auto finish = [&](const Meta &q, int bi) {
        float *buf = bufs + (size_t)(bi & 1) * nbatch * cstride;
        const int nb = min(nbatch, hi - lo - bi * nbatch);
        for (int slot = wave; slot < nb; slot += nw) {
            const int ks = __builtin_amdgcn_readlane(q.ks, slot);
            const int kf = ks & 0xffff, span = ks >> 16;
            if (span == 0) continue;  // no row reaches it: never read
            float *cyl = buf + (size_t)slot * cstride;
            const int nch = (span + kWave - 1) / kWave;
            if (GEOM == MBX_GEOM_CONE) {
                const float c0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, q.u0), slot));
                const float c1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, q.du), slot));
                const float u0 = fmaf(c1, (float)(kf + lane - g.kref), c0);  // v/sdd of this lane's slice in chunk 0
                const float du = c1 * (float)kWave;
                float *px = cyl + P + lane;
                // full chunks four at a time: the four LDS reads are in flight together
                float u0q[4];  // v/sdd of this lane's slice in chunk u (of each group of four)
#pragma unroll
                for (int u = 0; u < 4; ++u) u0q[u] = fmaf(du, (float)u, u0);
                int c = 0;
                for (; c + 4 <= nch - 1; c += 4) {
                    float x[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[u] = px[(c + u) * kWave];
                    const float cf = (float)c;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float uu = fmaf(du, cf, u0q[u]);
                        px[(c + u) * kWave] = x[u] * __builtin_amdgcn_sqrtf(fmaf(uu, uu, 1.0f));
                    }
                }
                for (; c < nch - 1; ++c) {
                    const float uu = fmaf(du, (float)c, u0);
                    px[c * kWave] = px[c * kWave] * __builtin_amdgcn_sqrtf(fmaf(uu, uu, 1.0f));
                }
                {  // last chunk: slices past the span read as zero
                    const int kk = (nch - 1) * kWave + lane;
                    const float uu = fmaf(du, (float)(nch - 1), u0);
                    const float x = cyl[P + kk] * __builtin_amdgcn_sqrtf(fmaf(uu, uu, 1.0f));
                    cyl[P + kk] = kk < span ? x : 0.0f;
                }
            } else {
                for (int c = 0; c < nch; ++c) {
                    const int kk = c * kWave + lane;
                    if (kk >= span) cyl[P + kk] = 0.0f;
                }
            }
            if (lane < P) {
                cyl[lane] = 0.0f;
                cyl[P + span + lane] = 0.0f;
            }
        }
    };

    float mrow[RPL];  // this lane's rows
#pragma unroll
    for (int r = 0; r < RPL; ++r) mrow[r] = (float)(wm0 + lane + kWave * r);
    const int wgrp = (wm0 >> 7) & 31;  // this wave's first 128-row group in the record masks
    constexpr unsigned gbits = (1u << (RPL / 2)) - 1;  // and its groups (64 * RPL rows)

    // vertical resampling of candidate record i (slot s of the batch at LDS base
    // sb) onto this lane's rows
    auto vert = [&](const float4 &R0, const float4 &R1, const float4 &R2, const float4 &R3, int i, int slot, int sb,
                    float (&v)[RPL]) {
        if (GEOM == MBX_GEOM_CONE) {
            const float W = R0.w, Wlo = R1.x, acr = R1.y, invW = R1.z, nai = R1.w, aif = R2.x;
            const float clo = R2.y, chi = R2.z, q1 = R3.y;
            // base at the lowest tap (LDS offsets are unsigned)
            int ib = sb - (__float_as_int(R2.w) & 0xffff);
            asm("" : "+s"(ib));  // one scalar base: the address is one shift-add per row
            if (VM == kVertTwo) {
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    const float mp = mrow[r] - aif;  // m - ai, exact
                    const float jf = floorf(__builtin_amdgcn_fmed3f(fmaf(mp, invW, nai), clo, chi));
                    // t_k0 - m - (1-W)/2: the first fma rounds once, at |t - m| < W
                    const float dm = fmaf(W, jf, -mp) + fmaf(Wlo, jf, acr);
                    const float f = __builtin_amdgcn_fmed3f(-dm, 0.0f, 1.0f);
                    const float *cp = lds + (ib + (int)jf);
                    const float c0 = cp[0], c1 = cp[1];
                    v[r] = fmaf(f, c1 - c0, c0);
                }
            } else if (VM == kVertUnit) {
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    const float mp = mrow[r] - aif;  // m - ai, exact
                    const float kr = rintf(__builtin_amdgcn_fmed3f(fmaf(mp, invW, nai), clo, chi));
                    // t - m for slice kref + kr: the first fma rounds once, at |t - m| < 1
                    const float dm = fmaf(W, kr, -mp) + fmaf(Wlo, kr, acr);
                    const float *cp = lds + (ib + (int)kr);  // slice kr - 1
                    const float cm = cp[0], c0 = cp[1], cq = cp[2];
                    // |dm| <= W/2, so |t -+ W - m| = W -+ dm
                    const float wm = __builtin_amdgcn_fmed3f(q1 + dm, 0.0f, 1.0f);
                    const float wq = __builtin_amdgcn_fmed3f(q1 - dm, 0.0f, 1.0f);
                    v[r] = fmaf(wq, cq - c0, fmaf(wm, cm - c0, c0));
                }
            } else {
                const float *rec = Tv + (size_t)lv[i] * kRec;
                const float hv = rec[9], Lv = rec[10], q2 = q1 - W;  // q2: taps two slices off the nearest
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    const float mp = mrow[r] - aif;  // m - ai, exact
                    const float kr = rintf(__builtin_amdgcn_fmed3f(fmaf(mp, invW, nai), clo, chi));
                    const float dm = fmaf(W, kr, -mp) + fmaf(Wlo, kr, acr);
                    const float *cp = lds + (ib + (int)kr) + RB;  // slice kr
                    // |dm| <= W/2, so |dm -+ jW| = jW -+ dm
                    float x = __builtin_amdgcn_fmed3f(hv - fabsf(dm), 0.0f, Lv) * cp[0];
                    x = fmaf(__builtin_amdgcn_fmed3f(q1 + dm, 0.0f, Lv), cp[-1], x);
                    x = fmaf(__builtin_amdgcn_fmed3f(q1 - dm, 0.0f, Lv), cp[1], x);
                    if (RB == 2) {
                        x = fmaf(__builtin_amdgcn_fmed3f(q2 + dm, 0.0f, Lv), cp[-2], x);
                        x = fmaf(__builtin_amdgcn_fmed3f(q2 - dm, 0.0f, Lv), cp[2], x);
                    }
                    v[r] = x;
                }
            }
        } else {
            // row m sees slice m (rows past the block read don't-cares)
            const int base = sb - (g.kref + P - RB) + P + wm0 - m0 + lane;  // slot's slice m0 at index P
#pragma unroll
            for (int r = 0; r < RPL; ++r) v[r] = lds[base + kWave * r];
        }
    };
    // bucket D is centre channel n0 + D - RAD: tap t lands on channel D - 2*RAD + t
    auto add_taps = [&](int D, const float4 &R0, int i, const float (&v)[RPL]) {
        if constexpr (SW) return;  // the window's add_window instead
        float Aall[5] = {0.0f, R0.x, R0.y, R0.z, 0.0f};  // Ac(n_c - 2 .. n_c + 2)
        if (RAD == 2) {
            const float *rec = Tv + (size_t)lv[i] * kRec;
            Aall[0] = rec[1];
            Aall[4] = rec[5];
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r)
#pragma unroll
            for (int t = 0; t <= 2 * RAD; ++t) {
                const int c = D - 2 * RAD + t;
                if (c >= 0 && c < CB) acc[r][c] = fmaf(Aall[2 - RAD + t], v[r], acc[r][c]);
            }
    };

    // window taps: tap t of the candidate lands on acc[.][t]
    auto add_window = [&](const float4 &R0, int i, const float (&v)[RPL]) {
        float Aall[5] = {0.0f, R0.x, R0.y, R0.z, 0.0f};  // Ac(n_c - 2 .. n_c + 2)
        if (RAD == 2) {
            const float *rec = Tv + (size_t)lv[i] * kRec;
            Aall[0] = rec[1];
            Aall[4] = rec[5];
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r)
#pragma unroll
            for (int t = 0; t <= 2 * RAD; ++t) acc[r][t] = fmaf(Aall[2 - RAD + t], v[r], acc[r][t]);
    };
    double ss = 0.0;  // this lane's share of sum w * (A delta)^2
    // store element (row m, channel n) of the output (optionally adding to what is there)
    auto put = [&](int m, int n, float v) {
        const size_t o = LAYOUT == MBX_LAYOUT_VCR ? ((size_t)vi * g.C + n) * (size_t)g.R + m
                                                  : ((size_t)vi * g.R + m) * (size_t)g.C + n;
        if (accumulate) v += to_float(sino[o]);
        const OT st = from_float<OT>(v);
        sino[o] = st;
        if (part_sumsq) {
            const float vs = to_float(st);
            ss += (double)vs * (double)vs * (w ? (double)w[o] : 1.0);
        }
    };
    // move the window to the next bucket: channel n0 - 2*RAD + dc is complete
    // (buckets up to its centre + RAD are added) and leaves it
    int dc = 0;
    int nextb = SW ? __builtin_amdgcn_readlane(bsl, 1) : 0;  // start of bucket dc + 1
    // (channel-major output) this lane's first row of channel n0 - 2*RAD + dc;
    // rows 64 apart are immediate offsets from it
    const bool all_rows = wm0 + kWave * RPL <= m1;  // wave-uniform
    OT *pch = sino + ((size_t)vi * g.C + (n0 - 2 * RAD)) * (size_t)g.R + wm0 + lane;
    const float *wch = w ? w + ((size_t)vi * g.C + (n0 - 2 * RAD)) * (size_t)g.R + wm0 + lane : nullptr;
    auto advance = [&]() {
        const int n = n0 - 2 * RAD + dc;
        if (n >= n0 && n < n1) {  // wave-uniform
            if (SW && all_rows) {
                float v[RPL];
#pragma unroll
                for (int r = 0; r < RPL; ++r) v[r] = acc[r][0];
                if (accumulate) {
#pragma unroll
                    for (int r = 0; r < RPL; ++r) v[r] += to_float(pch[kWave * r]);
                }
                OT st[RPL];
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    st[r] = from_float<OT>(v[r]);
                    pch[kWave * r] = st[r];
                }
                if (part_sumsq) {
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const double vs = (double)to_float(st[r]);
                        ss += vs * vs * (wch ? (double)wch[kWave * r] : 1.0);
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    const int m = wm0 + lane + kWave * r;
                    if (m < m1) put(m, n, acc[r][0]);
                }
            }
        }
        pch += g.R;
        if (wch) wch += g.R;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
#pragma unroll
            for (int t = 0; t < 2 * RAD; ++t) acc[r][t] = acc[r][t + 1];
            acc[r][2 * RAD] = 0.0f;
        }
        ++dc;
        nextb = __builtin_amdgcn_readlane(bsl, min(dc + 1, ND));
    };

    Meta mA{0, 0, 0.0f, 0.0f}, mB{0, 0, 0.0f, 0.0f};
    if (nbat > 0) {
        mA = load_meta(0);
        issue(mA, 0);
        mB = load_meta(min(1, nbat - 1));
    }
    for (int bi = 0; bi < nbat; ++bi) {
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of batch bi (and the next batch's data) have landed
        finish(mA, bi);
        __syncthreads();  // batch bi staged by every wave; batch bi-1 no longer read
        issue(mB, bi + 1);  // no slots past the last batch
        mA = mB;
        mB = load_meta(min(bi + 2, nbat - 1));
        const int b0 = lo + bi * nbatch;
        const int nb = min(nbatch, hi - b0);
        const int boff = (bi & 1) * nbatch * cstride;  // LDS float offset of this batch
        if constexpr (SW) {
            if (wm0 < m1) {  // wave-uniform
                // the batch's candidates in list order; the window follows their buckets
                const float4 *rp = (const float4 *)(FRv + (size_t)b0 * kFRec);
                int sb = boff + g.kref + P - (VM == kVertTwo ? 0 : RB);
                for (int slot = 0; slot < nb; ++slot, rp += 4, sb += cstride) {
#pragma clang loop unroll(disable)
                    while (b0 + slot >= nextb) advance();
                    const float4 R0 = rp[0], R1 = rp[1], R2 = rp[2], R3 = rp[3];
                    asm volatile("" ::"s"(R0.x), "s"(R0.y), "s"(R0.z), "s"(R0.w), "s"(R1.x), "s"(R1.y), "s"(R1.z),
                                 "s"(R1.w), "s"(R2.x), "s"(R2.y), "s"(R2.z), "s"(R2.w), "s"(R3.x), "s"(R3.y));
                    if (GEOM != MBX_GEOM_CONE || ((__float_as_uint(R3.x) >> wgrp) & gbits)) {
                        float v[RPL];
                        vert(R0, R1, R2, R3, b0 + slot, slot, sb, v);
                        add_window(R0, b0 + slot, v);
                    }
                }
            }
        } else if (wm0 < m1) {  // wave-uniform
#pragma unroll
            for (int D = 0; D < NDT; ++D) {
                const int slo = max(bs[D], b0) - b0;
                const int shi = min(bs[D + 1], b0 + nb) - b0;
                const float4 *rp = (const float4 *)(FRv + (size_t)(b0 + slo) * kFRec);
                // LDS index of slice kref of slot s, less its kf: cyl[i] is slice kf - P + i
                int sb = boff + slo * cstride + g.kref + P - (VM == kVertTwo ? 0 : RB);
                for (int slot = slo; slot < shi; ++slot, rp += 4, sb += cstride) {
                    const float4 R0 = rp[0], R1 = rp[1], R2 = rp[2], R3 = rp[3];
                    asm volatile("" ::"s"(R0.x), "s"(R0.y), "s"(R0.z), "s"(R0.w), "s"(R1.x), "s"(R1.y), "s"(R1.z),
                                 "s"(R1.w), "s"(R2.x), "s"(R2.y), "s"(R2.z), "s"(R2.w), "s"(R3.x), "s"(R3.y));
                    if (GEOM != MBX_GEOM_CONE || ((__float_as_uint(R3.x) >> wgrp) & gbits)) {
                        float v[RPL];
                        vert(R0, R1, R2, R3, b0 + slot, slot, sb, v);
                        add_taps(D, R0, b0 + slot, v);
                    }
                }
            }
        }
    }

    if constexpr (SW) {
        // the channels still in (or not yet reached by) the window
        if (wm0 < m1)
            while (dc < ND) advance();
    } else {
        // ---- write the register tile once
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int m = wm0 + lane + kWave * r;
            if (m < m1) {
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const int n = n0 + c;
                    if (n < n1) put(m, n, acc[r][c]);
                }
            }
        }
    }
    if (part_sumsq) {
        const double r = block_sum<double>(ss, dscratch);
        if (tid == 0) part_sumsq[b] = r;
    }
}

template <int GEOM, int RAD, int RB, int VM, typename OT, int LAYOUT>
static void launch_fwd_t(const DevGeom &g, const FwdPlan &pl, const float *T, const float *FR, const int *list,
                         const int *cstart, const float *vox, int64_t vstride, int npix, OT *sino, int accumulate,
                         const float *w, double *part, hipStream_t st) {
    // rows per lane of the window kernel: the plan's (the row-major kernel has its own)
    if constexpr (LAYOUT == MBX_LAYOUT_VCR) {
        if (pl.rpl == 2) {
            hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, VM, OT, LAYOUT, 2>), dim3((unsigned)pl.nblocks),
                               dim3(64 * pl.nw), pl.lds, st, g, T, FR, list, cstart, vox, vstride, npix, pl.ncb,
                               pl.nrb, pl.cb, pl.nb, pl.cstride, sino, accumulate, w, part);
            return;
        }
    }
    hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, VM, OT, LAYOUT, kFwdRPLW>), dim3((unsigned)pl.nblocks),
                       dim3(64 * pl.nw), pl.lds, st, g, T, FR, list, cstart, vox, vstride, npix, pl.ncb, pl.nrb, pl.cb,
                       pl.nb, pl.cstride, sino, accumulate, w, part);
}

template <int GEOM, int RAD, int RB, int VM>
static void dispatch_fwd_2(const DevGeom &g, const FwdPlan &pl, const float *T, const float *FR, const int *list,
                           const int *cstart, const float *vox, int64_t vstride, int npix, void *sino, int dt,
                           int layout, int accumulate, const float *w, double *part, hipStream_t st) {
    if (dt == MBX_DTYPE_F32) {
        if (layout == MBX_LAYOUT_VCR)
            launch_fwd_t<GEOM, RAD, RB, VM, float, MBX_LAYOUT_VCR>(g, pl, T, FR, list, cstart, vox, vstride, npix,
                                                                   (float *)sino, accumulate, w, part, st);
        else
            launch_fwd_t<GEOM, RAD, RB, VM, float, MBX_LAYOUT_VRC>(g, pl, T, FR, list, cstart, vox, vstride, npix,
                                                                   (float *)sino, accumulate, w, part, st);
    } else {
        if (layout == MBX_LAYOUT_VCR)
            launch_fwd_t<GEOM, RAD, RB, VM, bf16, MBX_LAYOUT_VCR>(g, pl, T, FR, list, cstart, vox, vstride, npix,
                                                                  (bf16 *)sino, accumulate, w, part, st);
        else
            launch_fwd_t<GEOM, RAD, RB, VM, bf16, MBX_LAYOUT_VRC>(g, pl, T, FR, list, cstart, vox, vstride, npix,
                                                                  (bf16 *)sino, accumulate, w, part, st);
    }
}

// Forward projection of pixels whose table is built.  aux holds
// fwd_aux_bytes(g, chunk >= npix, nviews) bytes; sumsq may be NULL.
int launch_fwd_tab(const DevGeom &g, const float *T, const int *NC, const float *vox, int npix, int nviews, void *sino,
                   int sino_dtype, int layout, int accumulate, const float *weights, void *aux, int chunk,
                   double *sumsq, hipStream_t st, int pstride, int steps) {
    const int64_t vstride = (int64_t)pstride * g.Nz;
    const FwdPlan pl = plan_fwd(g, nviews, layout);
    if (pl.lds > 160 * 1024)
        return fail(MBX_EUNSUPPORTED, "forward projection LDS footprint too large (num_recon_slices too big)");
    const int nbk = fwd_nbk(g);
    if (nbk > kBinMax) return fail(MBX_EUNSUPPORTED, "num_det_channels too large for the forward binning");
    if (!aux) return fail(MBX_EWORKSPACE, "forward projection needs a workspace");
    if (npix > chunk) return fail(MBX_EINVAL, "forward projection chunk larger than its workspace");
    char *a8 = (char *)aux;
    double *part = (double *)a8;
    int *list = (int *)(a8 + al256((size_t)fwd_part_blocks(g, nviews) * sizeof(double)));
    int *cstart = (int *)((char *)list + al256((size_t)nviews * chunk * sizeof(int)));
    float *FR = (float *)((char *)cstart + al256((size_t)nviews * (nbk + 1) * sizeof(int)));
    const int vm = fwd_vert_mode(g);
    if ((steps & kFwdBin) && nviews > 0) {
        // as many pass segments as the counters fit, up to one per wave
        const int nseg = std::min<int>(kBinWaves, (int)(kBinLds / sizeof(int)) / nbk);
        const size_t lds = (size_t)nseg * nbk * sizeof(int);
        hipLaunchKernelGGL(fwd_bin_kernel, dim3(nviews), dim3(kBinThreads), lds, st, g, T, NC, npix, nbk, vm, nseg,
                           list, cstart, FR);
        MBX_LAUNCH_CHECK();
    }
    if (!(steps & kFwdProject)) return MBX_OK;
    double *p = (sumsq || (steps & kFwdPartials)) ? part : nullptr;
#define MBX_FWD(G, RA, RB_, VM_)                                                                                      \
    dispatch_fwd_2<G, RA, RB_, VM_>(g, pl, T, FR, list, cstart, vox, vstride, npix, sino, sino_dtype, layout,        \
                                    accumulate, weights, p, st)
    if (is_fan(g)) {
        if (vm == kVertTwo) {
            if (g.rh == 1)
                MBX_FWD(MBX_GEOM_CONE, 1, 1, kVertTwo);
            else
                MBX_FWD(MBX_GEOM_CONE, 2, 1, kVertTwo);
        } else if (vm == kVertUnit) {
            if (g.rh == 1)
                MBX_FWD(MBX_GEOM_CONE, 1, 1, kVertUnit);
            else
                MBX_FWD(MBX_GEOM_CONE, 2, 1, kVertUnit);
        } else if (g.rh == 1 && g.rb == 1) {
            MBX_FWD(MBX_GEOM_CONE, 1, 1, kVertGeneral);
        } else if (g.rh == 1) {
            MBX_FWD(MBX_GEOM_CONE, 1, 2, kVertGeneral);
        } else if (g.rb == 1) {
            MBX_FWD(MBX_GEOM_CONE, 2, 1, kVertGeneral);
        } else {
            MBX_FWD(MBX_GEOM_CONE, 2, 2, kVertGeneral);
        }
    } else if (g.rh == 1) {
        MBX_FWD(MBX_GEOM_PARALLEL, 1, 1, kVertGeneral);
    } else {
        MBX_FWD(MBX_GEOM_PARALLEL, 2, 1, kVertGeneral);
    }
#undef MBX_FWD
    MBX_LAUNCH_CHECK();
    if (sumsq) return launch_sum_partials(part, pl.nblocks, sumsq, 0, st);
    return MBX_OK;
}

int launch_fwd(const DevGeom &g, const float *angles, const float *vox, const int32_t *pix, int npix,
               const int32_t *views, int nviews, void *sino, int sino_dtype, int layout, int accumulate,
               const float *weights, double *sumsq, void *ws, hipStream_t st) {
    if (nviews == 0) {
        if (sumsq) return launch_sum_partials(nullptr, 0, sumsq, 0, st);  // writes 0
        return MBX_OK;
    }
    if (!ws) return fail(MBX_EWORKSPACE, "forward projection needs a workspace");
    const int chunk = table_chunk(g, std::max(npix, 1), nviews);
    char *w8 = (char *)ws;
    float *T = (float *)w8;
    int *NC = table_nc(T, chunk, nviews);
    void *aux = w8 + table_bytes(chunk, nviews);
    if (npix == 0) {
        // nothing to project: the kernel still writes (zeros or the unchanged sinogram)
        return launch_fwd_tab(g, T, NC, vox, 0, nviews, sino, sino_dtype, layout, accumulate, weights, aux, chunk,
                              sumsq, st);
    }
    // Interleaved chunks (chunk c holds pixels c, c + nch, c + 2 nch, ...): a
    // run of whole image rows would put most of a view's candidates into a
    // few channel blocks, one chunk pixel spread does not.
    const int nch = (npix + chunk - 1) / chunk;
    for (int c = 0; c < nch; ++c) {
        const int n = (npix - c + nch - 1) / nch;
        int rc = launch_table(g, angles, pix + c, n, views, nviews, T, NC, st, nch);
        if (rc != MBX_OK) return rc;
        rc = launch_fwd_tab(g, T, NC, vox + (size_t)c * g.Nz, n, nviews, sino, sino_dtype, layout,
                            c == 0 ? accumulate : 1, weights, aux, chunk, c == nch - 1 ? sumsq : nullptr, st, nch);
        if (rc != MBX_OK) return rc;
    }
    return MBX_OK;
}

// ======================================================================
// Layout transpose (V,R,C) <-> (V,C,R), with dtype conversion
// ======================================================================

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
