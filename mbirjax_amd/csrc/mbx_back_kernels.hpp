// Back projector kernels and their launch chain (included by the translation
// units that instantiate dispatch_back_1 for one geometry each, so that the
// many kernel instantiations compile in parallel).  See mbx_back.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "mbx_internal.hpp"

namespace mbx {

// One wavefront per (pixel pair, slice chunk, view group), or per pixel for
// some waves (back_pairs).  For every view the wave reads the two pixels'
// records with scalar loads (one pixel-view ahead), gathers each pixel's
// weighted channel columns (2r+1, or the two with non-zero weight at radius
// 1) row-by-row into its LDS detector column (clamped, always-valid
// addresses; rows off the detector are zeroed afterwards), then every lane
// resamples both columns onto its slices.  Two pixels per wave keep all waves
// of a subset resident at once, so the waves of an XCD sweep the views
// together and each view is fetched from HBM about once.
// The 2r+1 channel columns of one pixel at one view: a buffer resource on the
// view (and on its weights), the byte offset of each column (channels off the
// detector are clamped to a valid column; their table weights are 0).
template <int RAD>
struct ColumnSource {
    i32x4_t rs, rw;
    int soff[2 * RAD + 1];
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
// TWO (UNIT with W <= 1): the window of half-width (W+1)/2 <= 1 reaches only
// rows floor(t) and floor(t)+1, at distances f and 1-f (f = t - floor(t)).
// Their weights clip((W+1)/2 - f, 0, W) and clip(f - (1-W)/2, 0, W) still sum
// to W, so the upper row carries med3(f - (1-W)/2, 0, W) and the lower row the
// rest.  The caller passes acd - 1/2, so that rint of the shifted position is
// floor(t) and f = 1/2 + d.
template <int KPL, int RAD, int POW, bool TAIL, bool FIT, bool UNIT, bool TWO = false>
__device__ __forceinline__ void back_vertical(const float *col, int shift, int n_rows, double acd, double Wd,
                                              float Wf, float hv, float Lv, float c0, float c1, float kr0,
                                              float krlast, float (&acc)[KPL]) {
    // Slice k sits at row ai + t with t = acr + W*(k - kref), formed in FP64 so
    // that the tap distances d = t - rint(t) carry no float32 rounding at the
    // magnitude of a row index; rint is the 1.5*2^52 shift, whose low word is
    // the rounded integer.
    constexpr double kMagic = 6755399441055744.0;
    const float *colp = col - shift;  // colp[m - RAD] is row ai + m (32-bit LDS address arithmetic)
    const float hm1 = hv - 1.0f, hW = 0.5f * Wf;
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
        if constexpr (TWO) {
            static_assert(RAD == 1 && UNIT, "the two-row form is a unit-sum radius-1 form");
            float clo, chi;  // rows floor(t), floor(t) + 1
            if (FIT) {
                const float *q = colp + mi;
                clo = q[1];
                chi = q[2];
            } else {
                const int base = mi - shift;
                clo = col[min(max(base + 1, 0), n_rows - 1)];
                chi = col[min(max(base + 2, 0), n_rows - 1)];
            }
            const float whi = __builtin_amdgcn_fmed3f(d + hW, 0.0f, Wf);
            sum = fmaf(whi, chi - clo, Wf * clo);
        } else if (RAD == 1) {
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

// One (view slot, pixel) record of the geometry table, read whole with one
// scalar load so that the next one can be in flight while this one is used.
typedef float BackRec __attribute__((ext_vector_type(kRec)));

// Marks every word of a record as read at this point.  A scalar load writes
// all 16 registers when it lands; words the kernel never reads would
// otherwise be reused while the load is in flight, and the write-after-write
// hazard costs a full wait right after the load is issued.
__device__ __forceinline__ void rec_touch(const BackRec &r) {
    asm volatile("" ::"s"(r[0]), "s"(r[1]), "s"(r[5]), "s"(r[11]), "s"(r[15]));
}

// Detector column col[i] = sum_j wc[j] * sinogram(column j, row mstart + i)
// for the NC channel columns at byte offsets soff[j] of the view (rs; rw the
// weights' view): whole 256-row groups 4 rows per lane (v4), the rest in
// 64-row chunks (loads of up to kBackFill chunks in flight).  Rows are clamped
// to the detector for the loads; the caller zeroes rows off it.
template <int NC, typename ET, bool WGT>
__device__ __forceinline__ void gather_column(const DevGeom &g, i32x4_t rs, i32x4_t rw, const int (&soff)[NC],
                                              const float (&wc)[NC], bool v4, int mstart, int n_rows, float *col,
                                              int lane) {
    const int ng = v4 ? n_rows >> 8 : 0;
    for (int gq = 0; gq < ng; ++gq) {
        const int r = min(max(mstart + gq * 4 * kWave + 4 * lane, 0), g.R - 4);
        mbx_f32x4 ev[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) ev[j] = buf_load4<ET>(rs, r * (int)sizeof(ET), soff[j]);
        mbx_f32x4 c;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float cq = 0.0f;
#pragma unroll
            for (int j = 0; j < NC; ++j) cq = fmaf(wc[j], ev[j][q], cq);
            c[q] = cq;
        }
        *(mbx_f32x4 *)(col + gq * 4 * kWave + 4 * lane) = c;  // 16-byte aligned
    }
    constexpr int FILL = kBackFill;
    const int nch = (n_rows + kWave - 1) / kWave;
    for (int c0 = 4 * ng; c0 < nch; c0 += FILL) {
        float ev[FILL][NC];
#pragma unroll
        for (int u = 0; u < FILL; ++u) {
            if (c0 + u < nch) {  // wave-uniform
                const int r = min(max(mstart + (c0 + u) * kWave + lane, 0), g.R - 1);
                const int vo = r * (int)sizeof(ET);
#pragma unroll
                for (int j = 0; j < NC; ++j) {
                    float e = buf_load<ET>(rs, vo, soff[j]);
                    if (WGT) e *= buf_load<float>(rw, r * 4, soff[j] * (int)(4 / sizeof(ET)));
                    ev[u][j] = e;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < FILL; ++u) {
            if (c0 + u < nch) {
                float c = 0.0f;
#pragma unroll
                for (int j = 0; j < NC; ++j) c = fmaf(wc[j], ev[u][j], c);
                col[(c0 + u) * kWave + lane] = c;  // cap holds whole chunks
            }
        }
    }
}

template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT, bool TAIL>
__device__ __forceinline__ void back_pixel_view(const DevGeom &g, const BackRec &rec,
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
        const int n = min(max(nc - RAD + j, 0), g.C - 1);
        src.soff[j] = n * g.R * (int)sizeof(ET);
    }
    if (GEOM == MBX_GEOM_CONE) {
        const float hv = rec[9], Lv = rec[10], c0 = rec[12], c1 = rec[13];
        const int ai = __float_as_int(rec[14]);
        const double Wd = (double)rec[7] + (double)rec[8], acd = (double)rec[6];
        // First and last rows the chunk's slices reach (a record of radius 1
        // holds them for the whole slice range)
        const bool whole = RAD == 1 && k0 == 0 && kend == g.Nz;  // wave-uniform
        const int mfirst = whole ? __float_as_int(rec[1])
                                 : ai + (int)rint(fma(Wd, (double)(k0 - g.kref), acd)) - RAD;
        const int mlast = whole ? __float_as_int(rec[5])
                                : ai + (int)rint(fma(Wd, (double)(kend - 1 - g.kref), acd)) + RAD;
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
        // ---- detector column col[i] = row mstart + i.  With radius-1 taps a
        // pixel rarely covers three channels: when an edge column has zero
        // weight (or lies off the detector) only the other two are gathered
        // (the same sums: the dropped term adds 0 * e).
        if (RAD == 1 && (Aw[0] == 0.0f || Aw[2 * RAD] == 0.0f)) {  // wave-uniform
            const int o = Aw[0] == 0.0f ? 1 : 0;
            const int so2[2] = {src.soff[o], src.soff[o + 1]};
            const float wc2[2] = {Aw[o], Aw[o + 1]};
            gather_column<2, ET, WGT>(g, src.rs, src.rw, so2, wc2, v4, mstart, n_rows, col, lane);
        } else {
            gather_column<2 * RAD + 1, ET, WGT>(g, src.rs, src.rw, src.soff, Aw, v4, mstart, n_rows, col, lane);
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
        if (CAN_UNIT && (flags & kBackUnit) && Wd <= 1.0) {  // wave-uniform
            if (fit)
                back_vertical<KPL, RAD, POW, TAIL, true, CAN_UNIT, CAN_UNIT>(col, shift, n_rows, acd - 0.5, Wd, Wf, hv,
                                                                             Lv, c0, c1, kr0, krl, acc);
            else
                back_vertical<KPL, RAD, POW, TAIL, false, CAN_UNIT, CAN_UNIT>(col, shift, n_rows, acd - 0.5, Wd, Wf,
                                                                              hv, Lv, c0, c1, kr0, krl, acc);
        } else if (CAN_UNIT && (flags & kBackUnit)) {
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
                c = fmaf(Aw[j], e, c);
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
    int accumulate, float *__restrict__ out, float *__restrict__ part, int npair) {
    constexpr int PPW = 2;  // pixels per wave
    extern __shared__ float lds[];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int split = blockIdx.x % nsplit;
    const int grp = blockIdx.x / nsplit;
    const int chunk = chunk0 + grp % nchunk;
    const int q = __builtin_amdgcn_readfirstlane((grp / nchunk) * 4 + wave);
    // waves below npair take two pixels, the rest one (see back_pairs)
    const int sA = q < npair ? PPW * q : PPW * npair + (q - npair);
    if (sA >= npix) return;  // whole wavefront leaves; no block barriers below
    const bool hasB = q < npair && sA + 1 < npix;
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
    // Records are loaded one pixel-view ahead: B's while A is projected, the
    // next view's A while B is (vb < nviews, so the first load is in range).
    const size_t vstride = (size_t)npix * kRec;
    const float *TA = T + (size_t)sA * kRec, *TB = T + (size_t)sB * kRec;
    BackRec recA = *(const BackRec *)(TA + vb * vstride);
    rec_touch(recA);  // landed before the loop
    for (int vi = vb; vi < ve; ++vi) {
        const int vd = views ? views[vi] : vi;
        const BackRec recB = *(const BackRec *)(TB + vi * vstride);
        back_pixel_view<GEOM, KPL, RAD, ET, POW, WGT, TAIL>(g, recA, sino, w, vd, k0, kend, cap, flags, colA, lane,
                                                            accA);
        rec_touch(recB);
        const BackRec recA1 = *(const BackRec *)(TA + min(vi + 1, ve - 1) * vstride);
        if (hasB)
            back_pixel_view<GEOM, KPL, RAD, ET, POW, WGT, TAIL>(g, recB, sino, w, vd, k0, kend, cap, flags, colB, lane,
                                                                accB);
        // The next A record has landed long since (B's projection waited on
        // the shared counter); waiting for it here, on every path, keeps the
        // use of the record at the loop top from also waiting for B's next
        // load (scalar loads return out of order: any wait is a full wait).
        rec_touch(recA1);
        recA = recA1;
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

// Waves of a view group that take two pixels (the others take one).  Two
// pixels per wave, unless the subset has more pixels than one XCD holds waves
// at once but fewer pairs, and the pairs leave at most half of the XCD's 128
// SIMDs one wave more than the rest (so those SIMDs set the time): then every
// wave slot of the XCD gets a wave (npix - slots waves take two pixels, the
// rest one).  One view group per XCD; 7 waves per SIMD by registers, fewer
// if the blocks' LDS columns do not fit 7 to a CU.  A one-pixel wave costs
// more per pixel than a pair, so balanced pairs (a multiple of 128) or
// pairs that load most SIMDs alike are kept.
inline int back_pairs(const BackPlan &pl, int npix, size_t lds_block) {
    const int pairs = (npix + 1) / 2;
    if (pl.nsplit != kXcds || lds_block == 0) return pairs;
    const int per_cu = (int)std::min<size_t>(7, kLdsPerCu / lds_block);  // 4-wave blocks per CU at once
    const int simds = kCusPerXcd * 4, slots = simds * per_cu;
    const int extra = pairs % simds;  // SIMDs holding one pair wave more than the others
    return (npix > slots && pairs < slots && extra > 0 && 2 * extra <= simds) ? npix - slots : pairs;
}

template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT, bool TAIL>
static void launch_back_range(const DevGeom &g, const BackPlan &pl, const float *T, const ET *sino, const float *w,
                              int npix, const int32_t *views, int nviews, int chunk0, int nch, float scale,
                              int accumulate, float *out, float *part, hipStream_t st) {
    constexpr int PPW = 2;  // pixels per wave, as in back_kernel
    const size_t lds = GEOM == MBX_GEOM_CONE ? 4 * PPW * (size_t)pl.cap * sizeof(float) : 0;
    const int npair = back_pairs(pl, npix, lds);
    const int64_t nwaves = npair + std::max(0, npix - PPW * npair);
    const int64_t groups = (nwaves + 3) / 4;  // 4 waves per block
    const int64_t blocks = groups * nch * pl.nsplit;
    const bool vec4 = g.R % 4 == 0 && ((uintptr_t)sino & (4 * sizeof(ET) - 1)) == 0;
    const int flags = (pl.unit ? kBackUnit : 0) | (vec4 ? kBackVec4 : 0);
    hipLaunchKernelGGL((back_kernel<GEOM, KPL, RAD, ET, POW, WGT, TAIL>), dim3((unsigned)blocks), dim3(256), lds, st,
                       g, T, sino, w, npix, views, nviews, pl.nsplit, chunk0, nch, pl.cap, flags, scale, accumulate,
                       out, part, npair);
}

// Full slice chunks in one launch, a partial last chunk in a second one.
template <int GEOM, int KPL, int RAD, typename ET, int POW, bool WGT>
static void launch_back_w(const DevGeom &g, const BackPlan &pl, const float *T, const ET *sino, const float *w,
                          int npix, const int32_t *views, int nviews, float scale, int accumulate, float *out,
                          float *part, hipStream_t st) {
    const int full = g.Nz / (kWave * KPL);
    if (full > 0)
        launch_back_range<GEOM, KPL, RAD, ET, POW, WGT, false>(g, pl, T, sino, w, npix, views, nviews, 0, full, scale,
                                                               accumulate, out, part, st);
    if (full < pl.nchunk)
        launch_back_range<GEOM, KPL, RAD, ET, POW, WGT, true>(g, pl, T, sino, w, npix, views, nviews, full, 1, scale,
                                                              accumulate, out, part, st);
}

template <int GEOM, int KPL, int RAD, typename ET, int POW>
static void launch_back_t(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, const float *w,
                          int npix, const int32_t *views, int nviews, float scale, int accumulate, float *out,
                          float *part, hipStream_t st) {
    const ET *se = (const ET *)sino;
    if (w)
        launch_back_w<GEOM, KPL, RAD, ET, POW, true>(g, pl, T, se, w, npix, views, nviews, scale, accumulate, out,
                                                     part, st);
    else
        launch_back_w<GEOM, KPL, RAD, ET, POW, false>(g, pl, T, se, w, npix, views, nviews, scale, accumulate, out,
                                                      part, st);
}

// Element types: float, bfloat16 or float16 sinograms (coeff_power 2: float only).
template <int GEOM, int KPL, int RAD>
static int dispatch_back_3(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, int dt,
                           const float *w, int npix, const int32_t *views, int nviews, int pw, float scale,
                           int accumulate, float *out, float *part, hipStream_t st) {
    if (pw == 2) {
        if (dt != MBX_DTYPE_F32) return fail(MBX_EUNSUPPORTED, "coeff_power=2 supports a float32 sinogram only");
        launch_back_t<GEOM, KPL, RAD, float, 2>(g, pl, T, sino, w, npix, views, nviews, scale, accumulate, out, part,
                                                st);
    } else if (dt == MBX_DTYPE_F32) {
        launch_back_t<GEOM, KPL, RAD, float, 1>(g, pl, T, sino, w, npix, views, nviews, scale, accumulate, out, part,
                                                st);
    } else if (dt == MBX_DTYPE_BF16) {
        launch_back_t<GEOM, KPL, RAD, bf16, 1>(g, pl, T, sino, w, npix, views, nviews, scale, accumulate, out, part,
                                               st);
    } else {
        launch_back_t<GEOM, KPL, RAD, f16, 1>(g, pl, T, sino, w, npix, views, nviews, scale, accumulate, out, part,
                                              st);
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
int dispatch_back_1(const DevGeom &g, const BackPlan &pl, const float *T, const void *sino, int dt, const float *w,
                    int npix, const int32_t *views, int nviews, int pw, float scale, int accumulate, float *out,
                    float *part, hipStream_t st) {
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

}  // namespace mbx
