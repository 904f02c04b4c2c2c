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
    if (g->kind != MBX_GEOM_PARALLEL && g->kind != MBX_GEOM_CONE)
        return fail(MBX_EINVAL, "geometry kind must be MBX_GEOM_PARALLEL or MBX_GEOM_CONE");
    if (g->num_views <= 0 || g->num_det_rows <= 0 || g->num_det_channels <= 0 || g->num_recon_rows <= 0 ||
        g->num_recon_cols <= 0 || g->num_recon_slices <= 0)
        return fail(MBX_EINVAL, "all sinogram and recon dimensions must be positive");
    if ((int64_t)g->num_recon_rows * g->num_recon_cols > INT32_MAX)
        return fail(MBX_EINVAL, "num_recon_rows*num_recon_cols exceeds int32");
    if (!(g->delta_det_channel > 0) || !(g->delta_voxel > 0) || !std::isfinite(g->delta_det_channel) ||
        !std::isfinite(g->delta_voxel))
        return fail(MBX_EINVAL, "delta_det_channel and delta_voxel must be positive and finite");
    if (g->psf_radius < 1 || g->psf_radius > 2)
        return fail(MBX_EUNSUPPORTED, "psf_radius must be 1 or 2 (got " + std::to_string(g->psf_radius) + ")");
    if (g->kind == MBX_GEOM_CONE) {
        if (!(g->delta_det_row > 0) || !std::isfinite(g->delta_det_row))
            return fail(MBX_EINVAL, "delta_det_row must be positive and finite");
        if (!(g->magnification > 0) || !std::isfinite(g->magnification))
            return fail(MBX_EINVAL, "magnification must be positive and finite");
        if (!(g->source_detector_dist > 0)) return fail(MBX_EINVAL, "source_detector_dist must be positive");
        if (g->bp_psf_radius < 1 || g->bp_psf_radius > 2)
            return fail(MBX_EUNSUPPORTED,
                        "bp_psf_radius must be 1 or 2 (got " + std::to_string(g->bp_psf_radius) + ")");
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
    o.rb = g->kind == MBX_GEOM_CONE ? g->bp_psf_radius : 1;
    o.dr = (float)(g->kind == MBX_GEOM_CONE ? g->delta_det_row : 1.0);
    o.dc = (float)g->delta_det_channel;
    o.row_off = (float)g->det_row_offset;
    o.ch_off = (float)g->det_channel_offset;
    o.inv_sdd = g->kind == MBX_GEOM_CONE ? (float)(1.0 / g->source_detector_dist) : 0.0f;
    o.inv_M = g->kind == MBX_GEOM_CONE ? (float)(1.0 / g->magnification) : 1.0f;
    o.dv = (float)g->delta_voxel;
    o.slice_off = (float)g->recon_slice_offset;
    o.row_center = (float)((g->num_det_rows - 1) / 2.0);
    o.ch_center = (float)((g->num_det_channels - 1) / 2.0);
    o.img_row_center = (float)((g->num_recon_rows - 1) / 2.0);
    o.img_col_center = (float)((g->num_recon_cols - 1) / 2.0);
    o.slice_center = (float)((g->num_recon_slices - 1) / 2.0);
    o.dv_over_dc = (float)(g->delta_voxel / g->delta_det_channel);
    o.dv_over_dr = g->kind == MBX_GEOM_CONE ? (float)(g->delta_voxel / g->delta_det_row) : 1.0f;
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
    const int64_t waves = (int64_t)npix * p.nchunk;
    // Small subsets: split the views over the 8 XCDs so that each XCD's L2
    // streams only its eighth of the sinogram; large passes are VALU-bound.
    p.nsplit = (waves <= 32768 && nviews >= 8) ? 8 : 1;
    return p;
}

int64_t back_workspace_bytes(const DevGeom &g, int npix, int nviews) {
    BackPlan p = plan_back(g, npix, nviews);
    if (p.nsplit == 1) return 0;
    return (int64_t)p.nsplit * npix * g.Nz * (int64_t)sizeof(float);
}

template <int GEOM, int KPL, int RAD, typename ET, int POW, bool KAHAN>
__global__ __launch_bounds__(256) void back_kernel(DevGeom g, const float *__restrict__ angles,
                                                   const ET *__restrict__ sino, const float *__restrict__ w,
                                                   const int32_t *__restrict__ pix, int npix,
                                                   const int32_t *__restrict__ views, int nviews, int nsplit,
                                                   int nchunk, float scale, int accumulate, float *__restrict__ out,
                                                   float *__restrict__ part) {
    extern __shared__ float lds[];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int split = blockIdx.x % nsplit;
    const int64_t grp = blockIdx.x / nsplit;
    const int chunk = (int)(grp % nchunk);
    const int64_t s = (grp / nchunk) * 4 + wave;
    if (s >= npix) return;  // whole wavefront leaves; no block barriers below
    const int p = pix[s];
    const int k0 = chunk * kWave * KPL;
    const int kend = min(g.Nz, k0 + kWave * KPL);
    const int vb = (int)(((int64_t)nviews * split) / nsplit);
    const int ve = (int)(((int64_t)nviews * (split + 1)) / nsplit);
    float *col = lds + wave * (g.R + 4);

    float acc[KPL], cmp[KPL];
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        acc[i] = 0.0f;
        cmp[i] = 0.0f;
    }

    for (int vi = vb; vi < ve; ++vi) {
        const int vd = views ? views[vi] : vi;
        float sn, cs;
        sincosf(angles[vd], &sn, &cs);
        const PV q = pv_full<GEOM>(g, pv_base<GEOM>(g, p, cs, sn));
        const int nc = (int)rintf(q.np);
        float Aw[2 * RAD + 1];
#pragma unroll
        for (int j = -RAD; j <= RAD; ++j) {
            float a = chan_weight(q, nc + j, g.C);
            Aw[j + RAD] = (POW == 2) ? a * a : a;
        }
        const size_t vbase = (size_t)vd * g.C;

        if (GEOM == MBX_GEOM_CONE) {
            const float t0 = fmaf(q.W, (float)k0, q.a);
            const float t1 = fmaf(q.W, (float)(kend - 1), q.a);
            const int mlo = max(0, (int)rintf(t0) - RAD);
            const int mhi = min(g.R - 1, (int)rintf(t1) + RAD);
            if (mlo > mhi) continue;  // wave-uniform: this view misses the detector
            const int span = mhi - mlo + 1;
            // detector column of this pixel: col[m] = sum_j Ac_j * e[v, nc+j, m]
            for (int m = lane; m < span; m += kWave) {
                float c = 0.0f;
#pragma unroll
                for (int j = -RAD; j <= RAD; ++j) {
                    const int n = nc + j;
                    if (n >= 0 && n < g.C) {
                        const size_t off = (vbase + n) * (size_t)g.R + (size_t)(mlo + m);
                        float e = to_float(sino[off]);
                        if (w) e *= w[off];
                        c = fmaf(Aw[j + RAD], e, c);
                    }
                }
                col[1 + m] = c;
            }
            if (lane == 0) {
                col[0] = 0.0f;
                col[span + 1] = 0.0f;
            }
            wave_sync();
#pragma unroll
            for (int i = 0; i < KPL; ++i) {
                const int k = k0 + lane + kWave * i;
                const float t = fmaf(q.W, (float)k, q.a);
                const float mc = rintf(t);
                const float d = t - mc;
                const int base = (int)mc - mlo + 1;
                float sum = 0.0f;
#pragma unroll
                for (int j = -RAD; j <= RAD; ++j) {
                    float wj = row_weight(q, d - (float)j);
                    if (POW == 2) wj *= wj;
                    const int idx = min(max(base + j, 0), span + 1);
                    sum = fmaf(wj, col[idx], sum);
                }
                const float u = fmaf(q.c1, (float)k, q.c0);
                const float s2 = fmaf(u, u, 1.0f);
                const float f = (POW == 2) ? s2 : __builtin_amdgcn_sqrtf(s2);
                const float contrib = sum * f;
                if (KAHAN) {
                    const float y = contrib - cmp[i];
                    const float tt = acc[i] + y;
                    cmp[i] = (tt - acc[i]) - y;
                    acc[i] = tt;
                } else {
                    acc[i] += contrib;
                }
            }
            wave_sync();
        } else {
#pragma unroll
            for (int i = 0; i < KPL; ++i) {
                const int k = k0 + lane + kWave * i;
                if (k < kend) {
                    float c = 0.0f;
#pragma unroll
                    for (int j = -RAD; j <= RAD; ++j) {
                        const int n = nc + j;
                        if (n >= 0 && n < g.C) {
                            const size_t off = (vbase + n) * (size_t)g.R + (size_t)k;
                            float e = to_float(sino[off]);
                            if (w) e *= w[off];
                            c = fmaf(Aw[j + RAD], e, c);
                        }
                    }
                    if (KAHAN) {
                        const float y = c - cmp[i];
                        const float tt = acc[i] + y;
                        cmp[i] = (tt - acc[i]) - y;
                        acc[i] = tt;
                    } else {
                        acc[i] += c;
                    }
                }
            }
        }
    }

#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        const int k = k0 + lane + kWave * i;
        if (k < kend) {
            if (nsplit == 1) {
                const size_t o = (size_t)s * g.Nz + k;
                const float v = scale * acc[i];
                out[o] = accumulate ? out[o] + v : v;
            } else {
                part[((size_t)split * npix + s) * g.Nz + k] = acc[i];
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

template <int GEOM, int KPL, int RAD, typename ET, int POW, bool KAHAN>
static void launch_back_t(const DevGeom &g, const BackPlan &pl, const float *angles, const ET *sino, const float *w,
                          const int32_t *pix, int npix, const int32_t *views, int nviews, float scale, int accumulate,
                          float *out, float *part, hipStream_t st) {
    const int64_t quads = ((int64_t)npix + 3) / 4;
    const int64_t blocks = quads * pl.nchunk * pl.nsplit;
    const size_t lds = GEOM == MBX_GEOM_CONE ? 4 * (size_t)(g.R + 4) * sizeof(float) : 0;
    hipLaunchKernelGGL((back_kernel<GEOM, KPL, RAD, ET, POW, KAHAN>), dim3((unsigned)blocks), dim3(256), lds, st, g,
                       angles, sino, w, pix, npix, views, nviews, pl.nsplit, pl.nchunk, scale, accumulate, out, part);
}

template <int GEOM, int KPL, int RAD>
static int dispatch_back_3(const DevGeom &g, const BackPlan &pl, const float *angles, const void *sino, int dt,
                           const float *w, const int32_t *pix, int npix, const int32_t *views, int nviews, int pw,
                           int kahan, float scale, int accumulate, float *out, float *part, hipStream_t st) {
    if (pw == 2) {
        if (dt != MBX_DTYPE_F32 || kahan)
            return fail(MBX_EUNSUPPORTED, "coeff_power=2 supports float32 input without Kahan summation");
        launch_back_t<GEOM, KPL, RAD, float, 2, false>(g, pl, angles, (const float *)sino, w, pix, npix, views, nviews,
                                                       scale, accumulate, out, part, st);
    } else if (dt == MBX_DTYPE_F32) {
        if (kahan)
            launch_back_t<GEOM, KPL, RAD, float, 1, true>(g, pl, angles, (const float *)sino, w, pix, npix, views,
                                                          nviews, scale, accumulate, out, part, st);
        else
            launch_back_t<GEOM, KPL, RAD, float, 1, false>(g, pl, angles, (const float *)sino, w, pix, npix, views,
                                                           nviews, scale, accumulate, out, part, st);
    } else {
        if (kahan)
            launch_back_t<GEOM, KPL, RAD, bf16, 1, true>(g, pl, angles, (const bf16 *)sino, w, pix, npix, views,
                                                         nviews, scale, accumulate, out, part, st);
        else
            launch_back_t<GEOM, KPL, RAD, bf16, 1, false>(g, pl, angles, (const bf16 *)sino, w, pix, npix, views,
                                                          nviews, scale, accumulate, out, part, st);
    }
    return MBX_OK;
}

template <int GEOM, int KPL>
static int dispatch_back_2(const DevGeom &g, const BackPlan &pl, const float *angles, const void *sino, int dt,
                           const float *w, const int32_t *pix, int npix, const int32_t *views, int nviews, int pw,
                           int kahan, float scale, int accumulate, float *out, float *part, hipStream_t st) {
    if (g.rh == 1)
        return dispatch_back_3<GEOM, KPL, 1>(g, pl, angles, sino, dt, w, pix, npix, views, nviews, pw, kahan, scale,
                                             accumulate, out, part, st);
    return dispatch_back_3<GEOM, KPL, 2>(g, pl, angles, sino, dt, w, pix, npix, views, nviews, pw, kahan, scale,
                                         accumulate, out, part, st);
}

template <int GEOM>
static int dispatch_back_1(const DevGeom &g, const BackPlan &pl, const float *angles, const void *sino, int dt,
                           const float *w, const int32_t *pix, int npix, const int32_t *views, int nviews, int pw,
                           int kahan, float scale, int accumulate, float *out, float *part, hipStream_t st) {
    switch (pl.kpl) {
        case 2:
            return dispatch_back_2<GEOM, 2>(g, pl, angles, sino, dt, w, pix, npix, views, nviews, pw, kahan, scale,
                                            accumulate, out, part, st);
        case 4:
            return dispatch_back_2<GEOM, 4>(g, pl, angles, sino, dt, w, pix, npix, views, nviews, pw, kahan, scale,
                                            accumulate, out, part, st);
        case 8:
            return dispatch_back_2<GEOM, 8>(g, pl, angles, sino, dt, w, pix, npix, views, nviews, pw, kahan, scale,
                                            accumulate, out, part, st);
        default:
            return dispatch_back_2<GEOM, 16>(g, pl, angles, sino, dt, w, pix, npix, views, nviews, pw, kahan, scale,
                                             accumulate, out, part, st);
    }
}

int launch_back(const DevGeom &g, const float *angles, const void *sino, int sino_dtype, const float *weights,
                const int32_t *pix, int npix, const int32_t *views, int nviews, int coeff_power, int kahan, float scale,
                int accumulate, float *out, void *ws, hipStream_t st) {
    if (npix == 0) return MBX_OK;
    const BackPlan pl = plan_back(g, npix, nviews);
    float *part = pl.nsplit > 1 ? (float *)ws : nullptr;
    if (pl.nsplit > 1 && !part) return fail(MBX_EWORKSPACE, "back projection needs a workspace for this size");
    int rc;
    if (g.kind == MBX_GEOM_CONE)
        rc = dispatch_back_1<MBX_GEOM_CONE>(g, pl, angles, sino, sino_dtype, weights, pix, npix, views, nviews,
                                            coeff_power, kahan, scale, accumulate, out, part, st);
    else
        rc = dispatch_back_1<MBX_GEOM_PARALLEL>(g, pl, angles, sino, sino_dtype, weights, pix, npix, views, nviews,
                                                coeff_power, kahan, scale, accumulate, out, part, st);
    if (rc != MBX_OK) return rc;
    MBX_LAUNCH_CHECK();
    if (pl.nsplit > 1) {
        const int64_t n = (int64_t)npix * g.Nz;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(back_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, part, pl.nsplit, n, scale,
                           accumulate, out);
        MBX_LAUNCH_CHECK();
    }
    return MBX_OK;
}

// ======================================================================
// Forward projection
// ======================================================================

constexpr int kFwdWaves = 4;
constexpr int kFwdRowsPerBlock = kFwdWaves * kWave;

FwdPlan plan_fwd(const DevGeom &g, int nviews) {
    FwdPlan p;
    p.cb = g.C >= 32 ? 32 : (g.C >= 16 ? 16 : (g.C >= 8 ? 8 : 4));
    p.ncb = (g.C + p.cb - 1) / p.cb;
    p.rows_pb = kFwdRowsPerBlock;
    p.nrb = (g.R + p.rows_pb - 1) / p.rows_pb;
    p.list = 512;
    const size_t tile = (size_t)p.cb * (p.rows_pb + 1) * sizeof(float);
    const size_t list = (size_t)p.list * 5 * sizeof(float);
    const size_t cyl = g.kind == MBX_GEOM_CONE ? (size_t)kFwdWaves * (g.Nz + 4) * sizeof(float) : 0;
    const size_t scan = 16 * sizeof(int);
    p.lds = tile + list + cyl + scan;
    p.nblocks = (int64_t)nviews * p.ncb * p.nrb;
    return p;
}

int64_t fwd_workspace_bytes(const DevGeom &g, int nviews) { return plan_fwd(g, nviews).nblocks * (int64_t)sizeof(double); }

template <typename T>
__device__ __forceinline__ T block_sum_256(T v, T *scratch) {
    // wave reduction with shuffles, then across the 4 waves through LDS
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

template <int GEOM, int RAD, int RB, typename OT, int LAYOUT>
__global__ __launch_bounds__(256) void fwd_kernel(DevGeom g, const float *__restrict__ angles,
                                                  const float *__restrict__ vox, const int32_t *__restrict__ pix,
                                                  int npix, const int32_t *__restrict__ views, int cb, int ncb,
                                                  int nrb, int list_cap, OT *__restrict__ sino, int accumulate,
                                                  const float *__restrict__ w, double *__restrict__ part_sumsq) {
    extern __shared__ float lds[];
    const int rows_pb = kFwdRowsPerBlock;
    const int tstride = rows_pb + 1;
    float *tile = lds;
    int *l_s = (int *)(tile + (size_t)cb * tstride);
    float *l_mag = (float *)(l_s + list_cap);
    float *l_np = l_mag + list_cap;
    float *l_cosa = l_np + list_cap;
    float *l_ctp = l_cosa + list_cap;
    float *cyl_all = l_ctp + list_cap;
    int *scan = (int *)(cyl_all + (GEOM == MBX_GEOM_CONE ? kFwdWaves * (g.Nz + 4) : 0));

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int64_t b = blockIdx.x;
    const int vi = (int)(b / ((int64_t)ncb * nrb));
    const int rem = (int)(b % ((int64_t)ncb * nrb));
    const int cbi = rem / nrb, rbi = rem % nrb;
    const int n0 = cbi * cb, n1 = min(g.C, n0 + cb);
    const int m0 = rbi * rows_pb, m1 = min(g.R, m0 + rows_pb);
    const int va = views ? views[vi] : vi;
    float sn, cs;
    sincosf(angles[va], &sn, &cs);

    for (int i = tid; i < cb * tstride; i += blockDim.x) tile[i] = 0.0f;
    float *cyl = cyl_all + wave * (g.Nz + 4);
    const int wm0 = m0 + wave * kWave;
    const int wm1 = min(m1, wm0 + kWave);
    const int m = wm0 + lane;
    const bool live = m < wm1;
    __syncthreads();

    int count = 0;
    for (int base = 0; base < npix; base += blockDim.x) {
        const int s = base + tid;
        bool flag = false;
        PVBase pb;
        pb.mag = pb.np = pb.cosa = pb.ctp = 0.0f;
        if (s < npix) {
            pb = pv_base<GEOM>(g, pix[s], cs, sn);
            const int nc = (int)rintf(pb.np);
            flag = (nc + RAD >= n0) && (nc - RAD <= n1 - 1);
            if (GEOM == MBX_GEOM_CONE && flag) {
                const PV q = pv_full<GEOM>(g, pb);
                const float ta = q.a;
                const float tb = fmaf(q.W, (float)(g.Nz - 1), q.a);
                flag = ((int)rintf(tb) + RAD + 1 >= m0) && ((int)rintf(ta) - RAD - 1 <= m1 - 1);
            }
        }
        const unsigned long long bal = __ballot(flag);
        const int wcount = __popcll(bal);
        const int lpre = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) scan[wave] = wcount;
        __syncthreads();
        int woff = 0, total = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            const int c = scan[i];
            if (i < wave) woff += c;
            total += c;
        }
        if (flag) {
            const int idx = count + woff + lpre;
            l_s[idx] = s;
            l_mag[idx] = pb.mag;
            l_np[idx] = pb.np;
            l_cosa[idx] = pb.cosa;
            l_ctp[idx] = pb.ctp;
        }
        count += total;
        __syncthreads();
        if (count > list_cap - (int)blockDim.x || base + (int)blockDim.x >= npix) {
            // ---- accumulate the listed pixels into this wave's 64 rows
            if (wm0 < wm1) {
                for (int c = 0; c < count; ++c) {
                    PVBase pbc;
                    pbc.mag = l_mag[c];
                    pbc.np = l_np[c];
                    pbc.cosa = l_cosa[c];
                    pbc.ctp = l_ctp[c];
                    const int sc = l_s[c];
                    const PV q = pv_full<GEOM>(g, pbc);
                    const int nc = (int)rintf(q.np);
                    float colv = 0.0f;
                    if (GEOM == MBX_GEOM_CONE) {
                        // slices whose taps can reach rows [wm0, wm1)
                        int kf = (int)rintf(((float)wm0 - q.a) * q.invW) - RB;
                        int kl = (int)rintf(((float)(wm1 - 1) - q.a) * q.invW) + RB;
                        kf = max(kf, 0);
                        kl = min(kl, g.Nz - 1);
                        if (kf > kl) continue;  // wave-uniform
                        const int span = kl - kf + 1;
                        const float *xs = vox + (size_t)sc * g.Nz;
                        for (int kk = lane; kk < span; kk += kWave) {
                            const int k = kf + kk;
                            const float u = fmaf(q.c1, (float)k, q.c0);
                            cyl[1 + kk] = xs[k] * __builtin_amdgcn_sqrtf(fmaf(u, u, 1.0f));
                        }
                        if (lane == 0) {
                            cyl[0] = 0.0f;
                            cyl[span + 1] = 0.0f;
                        }
                        wave_sync();
                        if (live) {
                            const float kc = rintf(((float)m - q.a) * q.invW);
                            const float dm = fmaf(q.W, kc, q.a) - (float)m;
                            const int bidx = (int)kc - kf + 1;
#pragma unroll
                            for (int j = -RB; j <= RB; ++j) {
                                const float wt = row_weight(q, fmaf(q.W, (float)j, dm));
                                const int idx = min(max(bidx + j, 0), span + 1);
                                colv = fmaf(wt, cyl[idx], colv);
                            }
                        }
                        wave_sync();
                    } else {
                        if (live) colv = vox[(size_t)sc * g.Nz + m];
                    }
                    if (live) {
#pragma unroll
                        for (int j = -RAD; j <= RAD; ++j) {
                            const int n = nc + j;
                            if (n >= n0 && n < n1) {
                                const float a = chan_weight(q, n, g.C);
                                atomicAdd(&tile[(n - n0) * tstride + (m - m0)], a * colv);
                            }
                        }
                    }
                }
            }
            count = 0;
            __syncthreads();
        }
    }
    __syncthreads();

    // ---- write the tile once (optionally adding to what is there)
    const int ncol = n1 - n0, nrow = m1 - m0;
    double ss = 0.0;
    if (LAYOUT == MBX_LAYOUT_VCR) {
        for (int nn = 0; nn < ncol; ++nn) {
            for (int mm = tid; mm < nrow; mm += blockDim.x) {
                const size_t o = ((size_t)vi * g.C + (n0 + nn)) * (size_t)g.R + (m0 + mm);
                float v = tile[nn * tstride + mm];
                if (accumulate) v += to_float(sino[o]);
                const OT st = from_float<OT>(v);
                sino[o] = st;
                if (part_sumsq) {
                    const float vs = to_float(st);
                    ss += (double)vs * (double)vs * (w ? (double)w[o] : 1.0);
                }
            }
        }
    } else {
        for (int idx = tid; idx < nrow * cb; idx += blockDim.x) {
            const int mm = idx / cb, nn = idx % cb;
            if (nn < ncol) {
                const size_t o = ((size_t)vi * g.R + (m0 + mm)) * (size_t)g.C + (n0 + nn);
                float v = tile[nn * tstride + mm];
                if (accumulate) v += to_float(sino[o]);
                const OT st = from_float<OT>(v);
                sino[o] = st;
                if (part_sumsq) {
                    const float vs = to_float(st);
                    ss += (double)vs * (double)vs * (w ? (double)w[o] : 1.0);
                }
            }
        }
    }
    if (part_sumsq) {
        double *dscratch = (double *)(scan + 8);
        const double r = block_sum_256<double>(ss, dscratch);
        if (tid == 0) part_sumsq[b] = r;
    }
}

template <int GEOM, int RAD, int RB, typename OT, int LAYOUT>
static void launch_fwd_t(const DevGeom &g, const FwdPlan &pl, const float *angles, const float *vox,
                         const int32_t *pix, int npix, const int32_t *views, OT *sino, int accumulate, const float *w,
                         double *part, hipStream_t st) {
    hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, OT, LAYOUT>), dim3((unsigned)pl.nblocks), dim3(256), pl.lds, st, g,
                       angles, vox, pix, npix, views, pl.cb, pl.ncb, pl.nrb, pl.list, sino, accumulate, w, part);
}

template <int GEOM, int RAD, int RB>
static void dispatch_fwd_2(const DevGeom &g, const FwdPlan &pl, const float *angles, const float *vox,
                           const int32_t *pix, int npix, const int32_t *views, void *sino, int dt, int layout,
                           int accumulate, const float *w, double *part, hipStream_t st) {
    if (dt == MBX_DTYPE_F32) {
        if (layout == MBX_LAYOUT_VCR)
            launch_fwd_t<GEOM, RAD, RB, float, MBX_LAYOUT_VCR>(g, pl, angles, vox, pix, npix, views, (float *)sino,
                                                               accumulate, w, part, st);
        else
            launch_fwd_t<GEOM, RAD, RB, float, MBX_LAYOUT_VRC>(g, pl, angles, vox, pix, npix, views, (float *)sino,
                                                               accumulate, w, part, st);
    } else {
        if (layout == MBX_LAYOUT_VCR)
            launch_fwd_t<GEOM, RAD, RB, bf16, MBX_LAYOUT_VCR>(g, pl, angles, vox, pix, npix, views, (bf16 *)sino,
                                                              accumulate, w, part, st);
        else
            launch_fwd_t<GEOM, RAD, RB, bf16, MBX_LAYOUT_VRC>(g, pl, angles, vox, pix, npix, views, (bf16 *)sino,
                                                              accumulate, w, part, st);
    }
}

int launch_fwd(const DevGeom &g, const float *angles, const float *vox, const int32_t *pix, int npix,
               const int32_t *views, int nviews, void *sino, int sino_dtype, int layout, int accumulate,
               const float *weights, double *sumsq, void *ws, hipStream_t st) {
    if (nviews == 0) {
        if (sumsq) MBX_HIP_CHECK(hipMemsetAsync(sumsq, 0, sizeof(double), st));
        return MBX_OK;
    }
    const FwdPlan pl = plan_fwd(g, nviews);
    if (pl.lds > 160 * 1024)
        return fail(MBX_EUNSUPPORTED, "forward projection LDS footprint too large (num_recon_slices too big)");
    double *part = nullptr;
    if (sumsq) {
        if (!ws) return fail(MBX_EWORKSPACE, "forward projection with sumsq needs a workspace");
        part = (double *)ws;
    }
    if (g.kind == MBX_GEOM_CONE) {
        if (g.rh == 1 && g.rb == 1)
            dispatch_fwd_2<MBX_GEOM_CONE, 1, 1>(g, pl, angles, vox, pix, npix, views, sino, sino_dtype, layout,
                                                accumulate, weights, part, st);
        else if (g.rh == 1)
            dispatch_fwd_2<MBX_GEOM_CONE, 1, 2>(g, pl, angles, vox, pix, npix, views, sino, sino_dtype, layout,
                                                accumulate, weights, part, st);
        else if (g.rb == 1)
            dispatch_fwd_2<MBX_GEOM_CONE, 2, 1>(g, pl, angles, vox, pix, npix, views, sino, sino_dtype, layout,
                                                accumulate, weights, part, st);
        else
            dispatch_fwd_2<MBX_GEOM_CONE, 2, 2>(g, pl, angles, vox, pix, npix, views, sino, sino_dtype, layout,
                                                accumulate, weights, part, st);
    } else {
        if (g.rh == 1)
            dispatch_fwd_2<MBX_GEOM_PARALLEL, 1, 1>(g, pl, angles, vox, pix, npix, views, sino, sino_dtype, layout,
                                                    accumulate, weights, part, st);
        else
            dispatch_fwd_2<MBX_GEOM_PARALLEL, 2, 1>(g, pl, angles, vox, pix, npix, views, sino, sino_dtype, layout,
                                                    accumulate, weights, part, st);
    }
    MBX_LAUNCH_CHECK();
    if (sumsq) return launch_sum_partials(part, pl.nblocks, sumsq, 0, st);
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
    (void)npix;
    return fwd_workspace_bytes(d, nviews);
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
    if (!dtype_ok(sino_dtype)) return fail(MBX_EINVAL, "unknown sinogram dtype");
    if (sino_layout != MBX_LAYOUT_VCR && sino_layout != MBX_LAYOUT_VRC)
        return fail(MBX_EINVAL, "unknown sinogram layout");
    if (!angles || !sino || (npix > 0 && (!vox || !pix)))
        return fail(MBX_EINVAL, "NULL device pointer (angles, sino, vox or pix)");
    if (sumsq && workspace_bytes < fwd_workspace_bytes(d, nviews))
        return fail(MBX_EWORKSPACE, "forward workspace too small: need " +
                                        std::to_string(fwd_workspace_bytes(d, nviews)) + " bytes");
    return launch_fwd(d, angles, vox, pix, npix, views, nviews, sino, sino_dtype, sino_layout, accumulate, weights,
                      sumsq, workspace, (hipStream_t)stream);
}

int mbx_back_project(const mbx_geom *g, const float *angles, const void *sino, int32_t sino_dtype,
                     int32_t num_sino_views, const float *weights, const int32_t *pix, int32_t npix,
                     const int32_t *views, int32_t nviews, int32_t coeff_power, int32_t kahan, float scale,
                     int32_t accumulate, float *out, void *workspace, int64_t workspace_bytes, void *stream) {
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
    return launch_back(d, angles, sino, sino_dtype, weights, pix, npix, views, nviews, coeff_power, kahan, scale,
                       accumulate, out, workspace, (hipStream_t)stream);
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
    if (src_dtype == MBX_DTYPE_F32 && dst_dtype == MBX_DTYPE_F32)
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
