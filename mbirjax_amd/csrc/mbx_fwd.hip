// Sparse forward projector on gfx950 (Projectors.sparse_forward_project,
// mbirjax/projectors.py:61-129): per view, the subset pixels are binned by
// centre channel; one workgroup per (view, channel block, row block) stages
// its candidates' cylinders in LDS and resamples them onto its rows, so every
// output element has one writer and a fixed summation order (deterministic).
// The operator and the per-(pixel, view) table are described in mbx_project.hip.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <type_traits>

#include "mbx_internal.hpp"

namespace mbx {
// ======================================================================
// Forward projection
// ======================================================================

constexpr int kFwdCB = 16;        // channels per block, (views, rows, channels) output: the per-lane register tile
constexpr int kFwdCBW = 32;       // channels per block, (views, channels, rows) output: the sliding window's sweep
constexpr int kFwdCBWMin = 8;     // ... narrowed down to this on small launches
constexpr int64_t kFwdMinWaves = 16384;  // channel blocks narrow below this many waves (16 per SIMD)
constexpr int64_t kFwdFewWaves = 4096;   // four waves per SIMD: a "few-wave" launch (its staging batch)
constexpr int kFwdRPL = 2;        // detector rows per lane (64 apart), (views, rows, channels) output
constexpr int kFwdRPLW = 4;       // detector rows per lane, (views, channels, rows) output (2 on
                                  // detectors of at most 256 rows)
constexpr int kFwdPad = 8;        // zero slices either side of a staged cylinder (>= 2*bp_psf_radius + 1;
                                  // a multiple of 4 keeps every staged cylinder 16-byte aligned)
constexpr int kFwdMaxBatch = 16;  // cylinders staged per batch
constexpr size_t kFwdLds = 36 * 1024;   // both staging buffers of a block, (views, rows, channels) output
constexpr size_t kFwdLdsW = 20 * 1024;  // the same, (views, channels, rows) output: small batches, more blocks per CU
constexpr int kBinMax = 16384;    // channel buckets the binning kernel holds in LDS
constexpr int kFRecLds = 16;      // floats per forward record staged in LDS (window kernel)
constexpr size_t kFwdRecLdsBytes = 2 * kFwdMaxBatch * kFRecLds * sizeof(float);  // its static LDS array

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
    // Launches of few waves (at most two rounds of kFwdFewWaves): a smaller
    // staging batch if that lets every block be resident at once (160 KB of
    // LDS per CU, 1 KB of allocation slack per block), down to 4 slots.
    const size_t slot_bytes = 2 * (size_t)p.cstride * sizeof(float);
    if (p.nblocks * p.nw <= 2 * kFwdFewWaves) {
        const int64_t per_cu = (p.nblocks + 255) / 256;
        const size_t fit = (size_t)(160 * 1024 / per_cu) - 1024 - 64 * sizeof(int) -
                           (layout == MBX_LAYOUT_VCR ? kFwdRecLdsBytes : 0);
        if (fit < budget && fit >= 4 * slot_bytes) budget = fit;
    }
    p.nb = (int)std::max<size_t>(1, std::min<size_t>(kFwdMaxBatch, budget / slot_bytes));
    // fan geometries with spans of at most 512 slices and blocks of two or
    // more waves stage through registers (fwd_reg_staging), two slots per wave
    // at most (single-wave blocks keep their larger batches)
    if (layout == MBX_LAYOUT_VCR && is_fan(g) && p.nw >= 2 && p.cstride <= 2 * 4 * kWave + 2 * kFwdPad + 8)
        p.nb = std::min(p.nb, 2 * p.nw);
    p.lds = 2 * (size_t)p.nb * p.cstride * sizeof(float) + 64 * sizeof(int);
    return p;
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// The window kernel stages each batch through registers (two 256-slice chunks
// per slot, at most two slots per wave) when the batch and the spans fit
// (plan_fwd caps the batch at two slots per wave for fan geometries).
static bool fwd_reg_staging(const FwdPlan &p) {
    return p.nw >= 2 && p.nb <= 2 * p.nw && p.cstride <= 2 * 4 * kWave + 2 * kFwdPad + 8;
}
// FP64 partial slots of the squared norm: one per wave of either output layout
static int64_t fwd_part_blocks(const DevGeom &g, int nviews) {
    const FwdPlan a = plan_fwd(g, nviews, MBX_LAYOUT_VCR), b = plan_fwd(g, nviews, MBX_LAYOUT_VRC);
    return std::max(a.nblocks * a.nw, b.nblocks * b.nw);
}
static int fwd_nbk(const DevGeom &g) { return g.C + 2 * g.rh; }  // centre-channel buckets n_c + r
int64_t fwd_partials(const DevGeom &g, int nviews, int layout) {
    const FwdPlan p = plan_fwd(g, nviews, layout);
    return p.nblocks * p.nw;
}

// Forward record of one (view, pixel) pair, stored in the view's binned order
// (centre channel, then pixel): what the forward update reads per candidate
// pixel, so that a block walks its candidates through contiguous scalar loads.
//   [0..2] Ac(n_c-1 .. n_c+1)   [3] W   [4] W_lo
//   [5] acr (acr - (1-W)/2 in the two-slice form)   [6] 1/W   [7] -acr/W
//   [8] ai (float)   [9], [10] clamp range of the resampling slice index,
//   relative to kref   [11] kf | span << 16: the slices kf .. kf+span-1 reach
//   a detector row; bit 31 set: the two-slice form in a unit-sum launch
//   [12] bit g: a row of 128g .. 128g+127 is reached (all
//   bits when the detector has more than 4096 rows)   [13] (1-W)/2
//   [14] v/sdd at kref   [15] its step per slice
constexpr int kFRec = 16;

// Binning blocks per view: a launch aims for kBinBlocks blocks (one block per
// view leaves most CUs idle when a table chunk holds tens of thousands of
// pixels over a few hundred views), each part holding kBinMinSlots slots or
// more.  Non-decreasing in npix, so a workspace sized for a chunk holds the
// part counts of every smaller pixel set.
constexpr int kBinBlocks = 2048;
constexpr int kBinMinSlots = 1024;
static int bin_parts(int npix, int nviews) {
    if (nviews <= 0) return 1;
    return std::max(1, std::min((kBinBlocks + nviews - 1) / nviews, npix / kBinMinSlots));
}
// per (view, part, bucket) slot counts of a binning split into parts
static size_t bin_count_bytes(const DevGeom &g, int chunk, int nviews) {
    const int np = bin_parts(chunk, nviews);
    return np > 1 ? al256((size_t)nviews * np * fwd_nbk(g) * sizeof(int)) : 0;
}

// Forward scratch after the geometry table: per-wave FP64 partials of the
// squared norm, then per view the pixel list sorted by centre channel, the
// bucket starts and the forward records in that order, then the binning's
// per-part bucket counts.
int64_t fwd_aux_bytes(const DevGeom &g, int chunk, int nviews) {
    return (int64_t)(al256((size_t)fwd_part_blocks(g, nviews) * sizeof(double)) +
                     al256((size_t)nviews * chunk * sizeof(int)) +
                     al256((size_t)nviews * (fwd_nbk(g) + 1) * sizeof(int)) +
                     al256((size_t)nviews * chunk * kFRec * sizeof(float)) + bin_count_bytes(g, chunk, nviews));
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

constexpr int kBinThreads = 1024;
constexpr int kBinWaves = kBinThreads / 64;
constexpr size_t kBinLds = 64 * 1024;  // counters [nseg][nbk]

// Binning split into parts, first pass: block view * nparts + part counts the
// slots of its part per centre-channel bucket into gtot[view][part][bucket]
// (counts, so the atomics' order does not matter).
__global__ __launch_bounds__(kBinThreads) void fwd_bin_count_kernel(DevGeom g, const int *__restrict__ NC, int npix, int nbk,
                                                             int nparts, int *__restrict__ gtot) {
    extern __shared__ int cnt[];  // [nbk]
    const int v = blockIdx.x / nparts, part = blockIdx.x - v * nparts;
    const int *ncv = NC + (size_t)v * npix;
    for (int i = threadIdx.x; i < nbk; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const int pb = (int)((int64_t)npix * part / nparts), pe = (int)((int64_t)npix * (part + 1) / nparts);
    for (int s = pb + (int)threadIdx.x; s < pe; s += blockDim.x) {
        const int b = ncv[s] + g.rh;
        if (b >= 0 && b < nbk) atomicAdd(&cnt[b], 1);
    }
    __syncthreads();
    int *dst = gtot + (size_t)blockIdx.x * nbk;
    for (int i = threadIdx.x; i < nbk; i += blockDim.x) dst[i] = cnt[i];
}

// Per view: the subset pixels sorted by centre channel n_c (bucket n_c + r),
// ascending pixel slot within a bucket, the bucket starts, and the forward
// records in that order.  Pixels whose taps miss the detector are dropped.
// Stable counting sort.  A view's slots are split into nparts contiguous
// parts, one block each (blockIdx = view * nparts + part); with more than one
// part, fwd_bin_count_kernel has counted every part's slots per bucket first
// (gtot), so each block knows how many slots of a bucket the parts before it
// hold.  Within a block, wave s (s < nseg) owns the s-th of nseg contiguous
// ranges of the part's slots; pass 1 counts its pixels per bucket, a scan
// over (bucket, segment) gives every segment its cursor in each bucket, and
// pass 2 walks the range 64 slots at a time in order, sorting the round's
// (bucket, lane) keys across the wave and ranking each slot within its bucket,
// so the order -- and the forward summation order -- is ascending slot within
// each bucket whatever the atomics' timing; the position goes back to the
// slot's own lane, which loaded the slot's table record while the round was
// sorted and writes its forward record there.
__global__ __launch_bounds__(kBinThreads) void fwd_bin_kernel(DevGeom g, const float *__restrict__ T,
                                                              const int *__restrict__ NC, int npix, int nbk,
                                                              int vmode, int nseg, int nparts,
                                                              const int *__restrict__ gtot, int *__restrict__ list,
                                                              int *__restrict__ cstart, float *__restrict__ FR) {
    extern __shared__ int cnt[];  // [nseg][nbk]
    const int rad = g.rh;
    const int v = blockIdx.x / nparts, part = blockIdx.x - v * nparts, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int *ncv = NC + (size_t)v * npix;
    int *lg = list + (size_t)v * npix;
    int *cs = cstart + (size_t)v * (nbk + 1);
    for (int i = tid; i < nseg * nbk; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const int pb = (int)((int64_t)npix * part / nparts);
    const int pn = (int)((int64_t)npix * (part + 1) / nparts) - pb;
    const int sb = pb + (int)((int64_t)pn * wave / nseg), se = pb + (int)((int64_t)pn * (wave + 1) / nseg);
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
    for (int b = b0; b < b1; ++b) {  // the view's slots in bucket b
        if (nparts == 1)
            for (int q = 0; q < nseg; ++q) run += cnt[q * nbk + b];
        else
            for (int p = 0; p < nparts; ++p) run += gtot[((size_t)v * nparts + p) * nbk + b];
    }
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
        if (part == 0) cs[b] = base;
        int cur = base, tot = 0;
        if (nparts > 1) {  // the slots of the parts before this one come first
            for (int p = 0; p < nparts; ++p) {
                const int c = gtot[((size_t)v * nparts + p) * nbk + b];
                if (p < part) cur += c;
                tot += c;
            }
        }
        for (int q = 0; q < nseg; ++q) {  // segment q's cursor in bucket b
            const int c = cnt[q * nbk + b];
            cnt[q * nbk + b] = cur;
            cur += c;
        }
        base += nparts > 1 ? tot : cur - base;
    }
    if (part == 0 && tid == blockDim.x - 1) cs[nbk] = base;
    __syncthreads();
    // forward record of slot s at list position i
    auto write_record = [&](int i, const float4 &R0, const float4 &R1, const float4 &R2, const float4 &R3) {
        float o[kFRec];
        o[0] = R0.z;
        o[1] = R0.w;
        o[2] = R1.x;
        int rlo = 0, rhi = g.R - 1, kf = 0, span = 0;
        bool two = false;
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
            // In the unit-sum mode a candidate with W >= 1 takes the two-slice
            // form (its third tap always weighs 0: a window of width W + 1 holds
            // at most two slices W apart), flagged in bit 31 of word 11.
            two = vmode == kVertTwo || (vmode == kVertUnit && (W > 1.0f || (W == 1.0f && Wlo >= 0.0f)));
            o[3] = W;
            o[4] = Wlo;
            o[5] = two ? acr - q1 : acr;
            o[6] = invW;
            o[7] = -(acr * invW);
            o[8] = (float)ai;
            o[13] = q1;
            o[14] = R3.x;
            o[15] = R3.y;
            // staged slice i of the cylinder is slice kf - P + i; slices kf-P ..
            // kf-1 and kf+span .. kf+span+P-1 are zeros.  Clamping the nearest
            // (or lower) slice keeps every tap there; a clamped row only meets zeros.
            if (two) {
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
        o[11] = __int_as_float(kf | (span << 16) | (two && vmode == kVertUnit ? (int)0x80000000u : 0));
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
    if (wave < nseg && sb < se) {  // pass 2: place in slot order; each lane writes one slot's record
        // This lane's slot of a round: its bucket and table record are loaded
        // a round ahead (lanes past the range load a valid slot, unused).
        const float *Tv = T + (size_t)v * npix * kRec;
        auto load_round = [&](int r0, float4 &A, float4 &B, float4 &C, float4 &D, int &nc) {
            const int sc = min(r0 + lane, se - 1);
            const float4 *r4 = (const float4 *)(Tv + (size_t)sc * kRec);
            A = r4[0];
            B = r4[1];
            C = r4[2];
            D = r4[3];
            nc = ncv[sc];
        };
        float4 N0, N1, N2, N3;
        int nnc;
        load_round(sb, N0, N1, N2, N3, nnc);
        for (int s0 = sb; s0 < se; s0 += 64) {
            const int s = s0 + lane;
            const float4 R0 = N0, R1 = N1, R2 = N2, R3 = N3;
            const int b = s < se ? nnc + rad : -1;
            if (s0 + 64 < se) load_round(s0 + 64, N0, N1, N2, N3, nnc);
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
template <int GEOM, int RAD, int RB, int VM, typename OT, int LAYOUT, int RPLW, bool RS>
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
    // (window kernel) the batches' forward records: an LDS array of its own,
    // apart from the cylinders' asynchronous copies
    float *recs = nullptr;
    if constexpr (SW) {
        __shared__ float4 recs_s[kFwdRecLdsBytes / sizeof(float4)];  // [2][kFwdMaxBatch][16] floats
        recs = (float *)recs_s;
    }

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
        float4 rq;  // (window kernel) words 4 (l % 4) .. + 3 of the batch's record l / 4
    };
    auto load_meta = [&](int bi) {
        Meta q;
        const int i = min(lo + bi * nbatch + min(lane, nbatch - 1), hi - 1);
        const float *rf = FRv + (size_t)i * kFRec;
        q.ks = GEOM == MBX_GEOM_CONE ? __float_as_int(rf[11]) : (m0 | ((m1 - m0) << 16));
        q.sc = lv[i];
        q.u0 = rf[14];
        q.du = rf[15];
        if constexpr (SW) {
            const int ri = min(lo + bi * nbatch + (lane >> 2), hi - 1);
            q.rq = *(const float4 *)(FRv + (size_t)ri * kFRec + 4 * (lane & 3));
        }
        return q;
    };
    // (window kernel, fan geometries) register staging: each wave loads the
    // raw slices of its (at most two) slots of the next batch into registers,
    // 16 bytes per lane, while the current batch is in use, and scales and
    // stores them to LDS when that batch starts: no asynchronous LDS copies
    // and no LDS read-back.  Spans of at most 512 slices (two 256-slice chunks).
    // RS: the launch satisfies fwd_reg_staging() (at most two slots per wave,
    // spans of at most 512 slices); the kernel then has no asynchronous copies.
    static_assert(!RS || (SW && GEOM == MBX_GEOM_CONE), "register staging: window kernel, fan geometries");
    mbx_f32x4 sreg[2][2];  // [slot s of this wave][chunk]
    // global -> registers (register staging) or asynchronous global -> LDS
    // copy of the raw slices of this wave's slots
    auto issue = [&](const Meta &q, int bi) {
        float *buf = bufs + (size_t)(bi & 1) * nbatch * cstride;
        const int nb = min(nbatch, hi - lo - bi * nbatch);
        if constexpr (RS) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int slot = wave + t * nw;  // < 8: a valid lane to read
                const int ks = slot < nb ? __builtin_amdgcn_readlane(q.ks, slot) : 0;
                const int kf = ks & 0xffff, span = (ks >> 16) & 0x7fff;
                const int sc = __builtin_amdgcn_readlane(q.sc, slot);
                // slices past the cylinder end read 0 (num_records bounds the buffer)
                const i32x4_t rsrc = make_rsrc_n(vox + sc * vstride + kf, (g.Nz - kf) * 4);
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    if (c * 4 * kWave < span) sreg[t][c] = raw_load_f32x4(rsrc, (c * kWave + lane) * 16, 0, 0);
            }
            return;
        }
        for (int slot = wave; slot < nb; slot += nw) {
            const int ks = __builtin_amdgcn_readlane(q.ks, slot);
            const int kf = ks & 0xffff, span = (ks >> 16) & 0x7fff;
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
        if (SW && wave == nw - 1 && lane < 4 * nb)  // the batch's forward records, read by every wave
            *(float4 *)(recs + (size_t)(bi & 1) * kFwdMaxBatch * kFRecLds + 4 * lane) = q.rq;
        if constexpr (RS) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int slot = wave + t * nw;
                if (slot >= nb) break;
                const int ks = __builtin_amdgcn_readlane(q.ks, slot);
                const int kf = ks & 0xffff, span = (ks >> 16) & 0x7fff;
                if (span == 0) continue;  // no row reaches it: never read
                float *cyl = buf + (size_t)slot * cstride;
                const float c0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, q.u0), slot));
                const float c1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, q.du), slot));
                const float du = c1 * (float)kWave;
                // lane l holds slices 256 c + 4 l + j: v/sdd as in the copy path,
                // from the slice's position within its 64-slice chunk and the chunk
                float w4[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) w4[j] = fmaf(c1, (float)(kf + 4 * (lane & 15) + j - g.kref), c0);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if (c * 4 * kWave >= span) break;
                    const float cf = (float)(4 * c + (lane >> 4));
                    mbx_f32x4 y;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float uu = fmaf(du, cf, w4[j]);
                        y[j] = sreg[t][c][j] * __builtin_amdgcn_sqrtf(fmaf(uu, uu, 1.0f));
                    }
                    // only lanes holding slices of the span store, so the slot's stride
                    // is never exceeded however short the span; the up to 3 slices past
                    // the span in the last lane's group are overwritten by the zero pad
                    // stored below (P >= 3, the same wave's later LDS store), and no
                    // tap reads beyond the pad
                    static_assert(kFwdPad >= 3, "the pad must cover a 4-slice group's overshoot");
                    if (c * 4 * kWave + 4 * lane < span)
                        *(mbx_f32x4 *)(cyl + P + c * 4 * kWave + 4 * lane) = y;  // 16-byte aligned
                }
                if (lane < P) {
                    cyl[lane] = 0.0f;
                    cyl[P + span + lane] = 0.0f;
                }
            }
            return;
        }
        for (int slot = wave; slot < nb; slot += nw) {
            const int ks = __builtin_amdgcn_readlane(q.ks, slot);
            const int kf = ks & 0xffff, span = (ks >> 16) & 0x7fff;
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
            // one scalar base: the address is one shift-add per row
            const int ib = __builtin_amdgcn_readfirstlane(sb - (__float_as_int(R2.w) & 0xffff));
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
                if (__float_as_int(R2.w) < 0) {  // this candidate's W >= 1: the two-slice form
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const float mp = mrow[r] - aif;  // m - ai, exact
                        const float jf = floorf(__builtin_amdgcn_fmed3f(fmaf(mp, invW, nai), clo, chi));
                        const float dm = fmaf(W, jf, -mp) + fmaf(Wlo, jf, acr);
                        const float f = __builtin_amdgcn_fmed3f(-dm, 0.0f, 1.0f);
                        const float *cp = lds + (ib + (int)jf) + RB;  // slice jf (the base is one slice lower)
                        const float c0 = cp[0], c1 = cp[1];
                        v[r] = fmaf(f, c1 - c0, c0);
                    }
                    return;
                }
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
    const bool hasw = w != nullptr;  // kernel arguments: wave-uniform tests
    const bool want_ss = part_sumsq != nullptr;
    // store element (row m, channel n) of the output (optionally adding to what is there)
    auto put = [&](int m, int n, float v) {
        const size_t o = LAYOUT == MBX_LAYOUT_VCR ? ((size_t)vi * g.C + n) * (size_t)g.R + m
                                                  : ((size_t)vi * g.R + m) * (size_t)g.C + n;
        if (accumulate) v += to_float(sino[o]);
        const OT st = from_float<OT>(v);
        store_stream(&sino[o], st);  // written once: kept out of L2
        if (want_ss) {
            const float vs = to_float(st);
            ss += (double)vs * (double)vs * (hasw ? (double)w[o] : 1.0);
        }
    };
    // move the window to the next bucket: channel n0 - 2*RAD + dc is complete
    // (buckets up to its centre + RAD are added) and leaves it
    int dc = 0;
    int nextb = SW ? __builtin_amdgcn_readlane(bsl, 1) : 0;  // start of bucket dc + 1
    // (channel-major output) channel n0 - 2*RAD + dc, rows wm0.. of this wave:
    // a wave-uniform base, lanes and rows 64 apart as offsets from it
    const bool all_rows = wm0 + kWave * RPL <= m1;  // wave-uniform
    const size_t ch0 = ((size_t)vi * g.C + (n0 - 2 * RAD)) * (size_t)g.R + wm0;
    OT *pch = sino + ch0;
    const float *wch = w + (hasw ? ch0 : 0);
    auto advance = [&]() {
        const int n = n0 - 2 * RAD + dc;
        if (n >= n0 && n < n1) {  // wave-uniform
            if (SW && all_rows) {
                float v[RPL];
#pragma unroll
                for (int r = 0; r < RPL; ++r) v[r] = acc[r][0];
                if (accumulate) {
#pragma unroll
                    for (int r = 0; r < RPL; ++r) v[r] += to_float(pch[lane + kWave * r]);
                }
                OT st[RPL];
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    st[r] = from_float<OT>(v[r]);
                    store_stream(&pch[lane + kWave * r], st[r]);  // written once: kept out of L2
                }
                if (want_ss) {
                    if (hasw) {
#pragma unroll
                        for (int r = 0; r < RPL; ++r) {
                            const double vs = (double)to_float(st[r]);
                            ss += vs * vs * (double)wch[lane + kWave * r];
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < RPL; ++r) {
                            const double vs = (double)to_float(st[r]);
                            ss += vs * vs;
                        }
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
        if (hasw) wch += g.R;
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
        if constexpr (!RS)
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
                // the batch's candidates in list order; the window follows their
                // buckets: the candidates of the current bucket in an inner loop
                // (nothing but their taps touches the window there), the window's
                // advances between buckets
                // records staged in LDS with the batch: vector operands, read
                // as 16-byte broadcasts
                const float4 *rp = (const float4 *)(recs + (size_t)(bi & 1) * kFwdMaxBatch * kFRecLds);
                int sb = boff + g.kref + P - (VM == kVertTwo ? 0 : RB);
                int slot = 0;
                for (;;) {
                    const int send = min(nextb - b0, nb);  // end of the current bucket in this batch
                    for (; slot < send; ++slot, rp += 4, sb += cstride) {
                        const float4 R0 = rp[0], R1 = rp[1], R2 = rp[2], R3 = rp[3];
                        asm volatile("" ::"v"(R0.x), "v"(R0.y), "v"(R0.z), "v"(R0.w), "v"(R1.x), "v"(R1.y),
                                     "v"(R1.z), "v"(R1.w), "v"(R2.x), "v"(R2.y), "v"(R2.z), "v"(R2.w), "v"(R3.x),
                                     "v"(R3.y));  // the record in one LDS round trip
                        if (GEOM != MBX_GEOM_CONE || ((__float_as_uint(R3.x) >> wgrp) & gbits)) {
                            float v[RPL];
                            vert(R0, R1, R2, R3, b0 + slot, slot, sb, v);
                            add_window(R0, b0 + slot, v);
                        }
                    }
                    if (slot >= nb) break;  // the bucket goes on in the next batch
                    advance();              // the next candidate starts a later bucket
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
    if (part_sumsq) {  // one partial per wave (no block-wide reduction)
        double r = ss;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
        if (lane == 0) part_sumsq[(size_t)b * nw + wave] = r;
    }
}

template <int GEOM, int RAD, int RB, int VM, typename OT, int LAYOUT>
static void launch_fwd_t(const DevGeom &g, const FwdPlan &pl, const float *T, const float *FR, const int *list,
                         const int *cstart, const float *vox, int64_t vstride, int npix, OT *sino, int accumulate,
                         const float *w, double *part, hipStream_t st) {
    // rows per lane of the window kernel: the plan's (the row-major kernel has its own)
    if constexpr (LAYOUT == MBX_LAYOUT_VCR) {
        if (pl.rpl == 2) {
            if constexpr (GEOM == MBX_GEOM_CONE) {
                if (fwd_reg_staging(pl)) {
                    hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, VM, OT, LAYOUT, 2, true>), dim3((unsigned)pl.nblocks),
                                       dim3(64 * pl.nw), pl.lds, st, g, T, FR, list, cstart, vox, vstride, npix, pl.ncb,
                                       pl.nrb, pl.cb, pl.nb, pl.cstride, sino, accumulate, w, part);
                    return;
                }
            }
            hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, VM, OT, LAYOUT, 2, false>), dim3((unsigned)pl.nblocks),
                               dim3(64 * pl.nw), pl.lds, st, g, T, FR, list, cstart, vox, vstride, npix, pl.ncb,
                               pl.nrb, pl.cb, pl.nb, pl.cstride, sino, accumulate, w, part);
            return;
        }
        if constexpr (GEOM == MBX_GEOM_CONE) {
            if (fwd_reg_staging(pl)) {
                hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, VM, OT, LAYOUT, kFwdRPLW, true>),
                                   dim3((unsigned)pl.nblocks), dim3(64 * pl.nw), pl.lds, st, g, T, FR, list, cstart,
                                   vox, vstride, npix, pl.ncb, pl.nrb, pl.cb, pl.nb, pl.cstride, sino, accumulate, w,
                                   part);
                return;
            }
        }
    }
    hipLaunchKernelGGL((fwd_kernel<GEOM, RAD, RB, VM, OT, LAYOUT, kFwdRPLW, false>), dim3((unsigned)pl.nblocks),
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
        // parts per view; their counts follow the records (bin_count_bytes)
        const int np = bin_parts(npix, nviews);
        if (np > bin_parts(chunk, nviews)) return fail(MBX_EINVAL, "forward binning split beyond its workspace");
        int *gtot = nullptr;
        if (np > 1) {
            gtot = (int *)((char *)FR + al256((size_t)nviews * chunk * kFRec * sizeof(float)));
            hipLaunchKernelGGL(fwd_bin_count_kernel, dim3(nviews * np), dim3(kBinThreads), (size_t)nbk * sizeof(int),
                               st, g, NC, npix, nbk, np, gtot);
            MBX_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(fwd_bin_kernel, dim3(nviews * np), dim3(kBinThreads), lds, st, g, T, NC, npix, nbk, vm, nseg,
                           np, gtot, list, cstart, FR);
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
    if (sumsq) return launch_sum_partials(part, pl.nblocks * pl.nw, sumsq, 0, st);
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

}  // namespace mbx
