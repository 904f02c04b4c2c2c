// Sparse back projector on gfx950 (Projectors.sparse_back_project,
// mbirjax/projectors.py:131-209): one wavefront per (pixel pair, slice chunk,
// view group), some waves with one pixel where that fills an XCD's wave slots
// evenly; for every view the wave gathers each pixel's weighted channel
// columns (contiguous rows in the channel-major sinogram) into an LDS detector
// column, then every lane resamples it onto its slices.  View groups are
// mapped to XCDs (blockIdx % 8) so each XCD's L2 streams only its share of the
// sinogram; the group partials are reduced in a fixed order (deterministic).
// The operator and the per-(pixel, view) table are described in mbx_project.hip.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <type_traits>

#include "mbx_internal.hpp"

namespace mbx {
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

// Back projection launches for one geometry (instantiated in
// mbx_back_cone.hip and mbx_back_parallel.hip).
template <int GEOM>
int dispatch_back_1(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, int dt, const float *w,
                    int npix, const int32_t *views, int nviews, int pw, float scale, int accumulate, float *out,
                    float *part, hipStream_t st);
extern template int dispatch_back_1<MBX_GEOM_CONE>(const DevGeom &, const BackPlan &, const float *, const void *,
                                                   int, const float *, int, const int32_t *, int, int, float, int,
                                                   float *, float *, hipStream_t);
extern template int dispatch_back_1<MBX_GEOM_PARALLEL>(const DevGeom &, const BackPlan &, const float *,
                                                       const void *, int, const float *, int, const int32_t *, int,
                                                       int, float, int, float *, float *, hipStream_t);

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
        rc = dispatch_back_1<MBX_GEOM_PARALLEL>(g, pl, T, sino, sino_dtype, weights, npix, views, nviews,
                                                coeff_power, scale, accumulate, out, part, st);
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

}  // namespace mbx
