// Device-side building blocks shared by the libmbx kernels (gfx950 / CDNA4).
//
// The projector geometry restates ConeBeamModel.compute_horizontal_data,
// compute_vertical_data_single_pixel and recon_ijk_to_xyz / geometry_xyz_to_uv_mag
// (mbirjax/cone_beam.py:543-681) and ParallelBeamModel.compute_proj_data
// (mbirjax/parallel_beam.py:223-284) as one per-(pixel, view) record that both the
// forward and the back projector evaluate with identical float32 operations, so
// the two kernels apply exactly transposed operators.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbx.h"

namespace mbx {

constexpr int kWave = 64;

// ---------------------------------------------------------------- bfloat16
// Storage type for the mixed-precision residual sinogram: the upper 16 bits of
// an IEEE float32, rounded to nearest-even on store.
struct bf16 {
    uint16_t bits;
};

// ---------------------------------------------------------------- float16
// Storage type for the 2-byte error sinogram of precision 'fp16': IEEE binary16
// (11-bit significand, 8x finer than bfloat16), rounded to nearest-even on
// store; the host keeps |e| within its range by scaling the problem by a
// power of two.
struct f16 {
    uint16_t bits;
};

__device__ __forceinline__ float h2f(uint32_t b16) { return (float)__builtin_bit_cast(_Float16, (uint16_t)b16); }
__device__ __forceinline__ uint32_t f2h(float x) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x); }

__device__ __forceinline__ float to_float(float x) { return x; }
__device__ __forceinline__ float to_float(bf16 x) { return __uint_as_float(((uint32_t)x.bits) << 16); }
__device__ __forceinline__ float to_float(f16 x) { return h2f(x.bits); }

template <typename T>
__device__ __forceinline__ T from_float(float x);
template <>
__device__ __forceinline__ float from_float<float>(float x) {
    return x;
}
template <>
__device__ __forceinline__ bf16 from_float<bf16>(float x) {
    // round to nearest even (NaN stays a quiet NaN): one v_cvt_pk_bf16_f32 on gfx950
    bf16 r;
    r.bits = __builtin_bit_cast(uint16_t, (__bf16)x);
    return r;
}
template <>
__device__ __forceinline__ f16 from_float<f16>(float x) {
    f16 r;  // v_cvt_f16_f32: round to nearest even
    r.bits = (uint16_t)f2h(x);
    return r;
}

// Streaming store (nontemporal hint): for outputs written once and not read
// again by the same kernel, so that they do not evict from L2 what the kernel
// does read again (the forward projector's staged voxel cylinders).
__device__ __forceinline__ void store_stream(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_stream(bf16 *p, bf16 v) { __builtin_nontemporal_store(v.bits, &p->bits); }
__device__ __forceinline__ void store_stream(f16 *p, f16 v) { __builtin_nontemporal_store(v.bits, &p->bits); }

// Four consecutive elements through one 16-byte (float) or 8-byte (bf16) access.
__device__ __forceinline__ void load4(const float *p, float v[4]) {
    const float4 t = *(const float4 *)p;
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
}
__device__ __forceinline__ void load4(const bf16 *p, float v[4]) {
    const uint2 t = *(const uint2 *)p;
    v[0] = __uint_as_float(t.x << 16);
    v[1] = __uint_as_float(t.x & 0xffff0000u);
    v[2] = __uint_as_float(t.y << 16);
    v[3] = __uint_as_float(t.y & 0xffff0000u);
}
__device__ __forceinline__ void load4(const f16 *p, float v[4]) {
    const uint2 t = *(const uint2 *)p;
    v[0] = h2f(t.x & 0xffffu);
    v[1] = h2f(t.x >> 16);
    v[2] = h2f(t.y & 0xffffu);
    v[3] = h2f(t.y >> 16);
}
__device__ __forceinline__ void store4(float *p, const float v[4]) { *(float4 *)p = make_float4(v[0], v[1], v[2], v[3]); }
__device__ __forceinline__ void store4(f16 *p, const float v[4]) {
    uint2 t;
    t.x = f2h(v[0]) | (f2h(v[1]) << 16);
    t.y = f2h(v[2]) | (f2h(v[3]) << 16);
    *(uint2 *)p = t;
}
__device__ __forceinline__ void store4(bf16 *p, const float v[4]) {
    uint2 t;
    t.x = (uint32_t)from_float<bf16>(v[0]).bits | ((uint32_t)from_float<bf16>(v[1]).bits << 16);
    t.y = (uint32_t)from_float<bf16>(v[2]).bits | ((uint32_t)from_float<bf16>(v[3]).bits << 16);
    *(uint2 *)p = t;
}

// Streaming variants (nontemporal: data touched once per pass, larger than the caches).
typedef float mbx_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t mbx_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load4_nt(const float *p, float v[4]) {
    const mbx_f32x4 t = __builtin_nontemporal_load((const mbx_f32x4 *)p);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
}
__device__ __forceinline__ void load4_nt(const bf16 *p, float v[4]) {
    const mbx_u32x2 t = __builtin_nontemporal_load((const mbx_u32x2 *)p);
    v[0] = __uint_as_float(t.x << 16);
    v[1] = __uint_as_float(t.x & 0xffff0000u);
    v[2] = __uint_as_float(t.y << 16);
    v[3] = __uint_as_float(t.y & 0xffff0000u);
}
__device__ __forceinline__ void load4_nt(const f16 *p, float v[4]) {
    const mbx_u32x2 t = __builtin_nontemporal_load((const mbx_u32x2 *)p);
    v[0] = h2f(t.x & 0xffffu);
    v[1] = h2f(t.x >> 16);
    v[2] = h2f(t.y & 0xffffu);
    v[3] = h2f(t.y >> 16);
}
__device__ __forceinline__ void store4_nt(f16 *p, const float v[4]) {
    mbx_u32x2 t;
    t.x = f2h(v[0]) | (f2h(v[1]) << 16);
    t.y = f2h(v[2]) | (f2h(v[3]) << 16);
    __builtin_nontemporal_store(t, (mbx_u32x2 *)p);
}
__device__ __forceinline__ void store4_nt(float *p, const float v[4]) {
    mbx_f32x4 t;
    t.x = v[0];
    t.y = v[1];
    t.z = v[2];
    t.w = v[3];
    __builtin_nontemporal_store(t, (mbx_f32x4 *)p);
}
__device__ __forceinline__ void store4_nt(bf16 *p, const float v[4]) {
    mbx_u32x2 t;
    t.x = (uint32_t)from_float<bf16>(v[0]).bits | ((uint32_t)from_float<bf16>(v[1]).bits << 16);
    t.y = (uint32_t)from_float<bf16>(v[2]).bits | ((uint32_t)from_float<bf16>(v[3]).bits << 16);
    __builtin_nontemporal_store(t, (mbx_u32x2 *)p);
}

// ---------------------------------------------------------------- buffer loads
// Raw buffer loads through a wave-uniform resource (base address in SGPRs):
// one 32-bit per-lane byte offset plus an SGPR byte offset, no 64-bit
// per-lane address arithmetic.  Offsets stay below 4 GiB of the base.
typedef int i32x4_t __attribute__((ext_vector_type(4)));
__device__ float raw_load_f32(i32x4_t rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.f32");
__device__ short raw_load_i16(i32x4_t rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.i16");
__device__ mbx_f32x4 raw_load_f32x4(i32x4_t rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.v4f32");
__device__ mbx_u32x2 raw_load_u32x2(i32x4_t rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.v2i32");

// Four consecutive elements at a byte offset (float32: 16 bytes, bfloat16: 8 bytes).
template <typename T>
__device__ __forceinline__ mbx_f32x4 buf_load4(i32x4_t rsrc, int voffset, int soffset);
template <>
__device__ __forceinline__ mbx_f32x4 buf_load4<float>(i32x4_t rsrc, int voffset, int soffset) {
    return raw_load_f32x4(rsrc, voffset, soffset, 0);
}
template <>
__device__ __forceinline__ mbx_f32x4 buf_load4<bf16>(i32x4_t rsrc, int voffset, int soffset) {
    const mbx_u32x2 u = raw_load_u32x2(rsrc, voffset, soffset, 0);
    mbx_f32x4 r;
    r.x = __uint_as_float(u.x << 16);
    r.y = __uint_as_float(u.x & 0xffff0000u);
    r.z = __uint_as_float(u.y << 16);
    r.w = __uint_as_float(u.y & 0xffff0000u);
    return r;
}

template <>
__device__ __forceinline__ mbx_f32x4 buf_load4<f16>(i32x4_t rsrc, int voffset, int soffset) {
    const mbx_u32x2 u = raw_load_u32x2(rsrc, voffset, soffset, 0);
    mbx_f32x4 r;
    r.x = h2f(u.x & 0xffffu);
    r.y = h2f(u.x >> 16);
    r.z = h2f(u.y & 0xffffu);
    r.w = h2f(u.y >> 16);
    return r;
}

__device__ __forceinline__ i32x4_t make_rsrc(const void *p) {
    const uint64_t a = (uint64_t)p;
    i32x4_t r;
    r.x = __builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    r.y = __builtin_amdgcn_readfirstlane((int)((uint32_t)(a >> 32) & 0xffffu));  // stride 0
    r.z = (int)0xffffffffu;                                                        // num_records: no clamp
    r.w = 0x00020000;                                                              // raw dword access (gfx9 family)
    return r;
}

// The same with num_records = nbytes: loads at byte offsets past it return 0.
__device__ __forceinline__ i32x4_t make_rsrc_n(const void *p, int nbytes) {
    i32x4_t r = make_rsrc(p);
    r.z = nbytes;
    return r;
}

template <typename T>
__device__ __forceinline__ float buf_load(i32x4_t rsrc, int voffset, int soffset);
template <>
__device__ __forceinline__ float buf_load<float>(i32x4_t rsrc, int voffset, int soffset) {
    return raw_load_f32(rsrc, voffset, soffset, 0);
}
template <>
__device__ __forceinline__ float buf_load<bf16>(i32x4_t rsrc, int voffset, int soffset) {
    const unsigned short b = (unsigned short)raw_load_i16(rsrc, voffset, soffset, 0);
    return __uint_as_float(((uint32_t)b) << 16);
}
template <>
__device__ __forceinline__ float buf_load<f16>(i32x4_t rsrc, int voffset, int soffset) {
    return h2f((unsigned short)raw_load_i16(rsrc, voffset, soffset, 0));
}

// ---------------------------------------------------------------- wave sync
// LDS hand-off between lanes of one wavefront (rocPRIM's wave_barrier idiom).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// ---------------------------------------------------------------- geometry
// Float32 copy of mbx_geom with the derived centres precomputed on the host.
struct DevGeom {
    int kind, V, R, C, rows, cols, Nz, rh, rb;
    float dr, dc, row_off, ch_off, inv_sdd, inv_M, dv, slice_off;
    float row_center;    // (R-1)/2
    float ch_center;     // (C-1)/2
    float img_row_center;  // (rows-1)/2
    float img_col_center;  // (cols-1)/2
    float slice_center;  // (Nz-1)/2
    float dv_over_dc, dv_over_dr;
    int kref;   // reference slice (Nz/2) of the table's vertical parametrisation
    mbx_geom h; // the caller's double-precision geometry (table kernel)
};

// Sum of v over the block (blockDim.x a multiple of 64), in a fixed order;
// valid in thread 0.  scratch: one element of shared memory per wave.
template <typename T>
__device__ __forceinline__ T block_reduce(T v, T *scratch) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += scratch[i];
    __syncthreads();
    return r;
}

// Cone and translation geometries share the fan projectors (the vertical and
// horizontal fans); only the per-(pixel, view) table differs.
__host__ __device__ inline bool is_fan(const DevGeom &g) { return g.kind != MBX_GEOM_PARALLEL; }

// The view-dependent part of one (pixel, view) pair that both projectors
// store or recompute: the pixel magnification, the fractional channel n_p,
// cos(alpha_xy) and cos(theta_p).
struct PVBase {
    float mag, np, cosa, ctp;
};

// Pixel p at a view with cos/sin (cs, sn):  recon_ijk_to_xyz, the pixel
// magnification 1/(1/M - y/sdd), u = mag*x, n_p = (u + det_channel_offset)/dc +
// (C-1)/2 and, for the cone, theta_p = atan2(u, sdd) expressed through
// cos/sin so that cos(angle - theta_p) needs no atan2.
template <int GEOM>
__device__ __forceinline__ PVBase pv_base(const DevGeom &g, int p, float cs, float sn) {
    const int i = p / g.cols;
    const int j = p - i * g.cols;
    const float yt = g.dv * ((float)i - g.img_row_center);
    const float xt = g.dv * ((float)j - g.img_col_center);
    const float x = cs * xt - sn * yt;
    PVBase b;
    if (GEOM == MBX_GEOM_CONE) {
        const float y = sn * xt + cs * yt;
        const float mag = 1.0f / (g.inv_M - y * g.inv_sdd);
        const float u = mag * x;
        b.mag = mag;
        b.np = (u + g.ch_off) / g.dc + g.ch_center;
        const float us = u * g.inv_sdd;
        const float ctp = 1.0f / sqrtf(1.0f + us * us);  // cos(theta_p)
        const float stp = us * ctp;                      // sin(theta_p)
        const float cd = cs * ctp + sn * stp;            // cos(angle - theta_p)
        const float sd = sn * ctp - cs * stp;            // sin(angle - theta_p)
        b.cosa = fmaxf(fabsf(cd), fabsf(sd));
        b.ctp = ctp;
    } else {
        b.mag = 1.0f;
        b.np = (x + g.ch_off) / g.dc + g.ch_center;
        b.cosa = fmaxf(fabsf(cs), fabsf(sn));
        b.ctp = 1.0f;
    }
    return b;
}

// Everything the taps need, derived from PVBase.
struct PV {
    float np;    // fractional channel
    float hc;    // (W_c + 1)/2
    float Lc;    // min(1, W_c)
    float sc;    // delta_voxel / cos(alpha_xy)
    // cone vertical fan: row position of slice k is t_k = a + W*k
    float a, W, hv, Lv, invW;
    // v_k / sdd = c0 + c1*k, so 1/cos(phi_k) = sqrt(1 + (c0 + c1*k)^2)
    float c0, c1;
};

template <int GEOM>
__device__ __forceinline__ PV pv_full(const DevGeom &g, const PVBase &b) {
    PV q;
    q.np = b.np;
    float Wc;
    if (GEOM == MBX_GEOM_CONE) {
        Wc = b.mag * g.dv_over_dc * (b.cosa / b.ctp);
    } else {
        Wc = g.dv_over_dc * b.cosa;
    }
    q.hc = 0.5f * (Wc + 1.0f);
    q.Lc = fminf(1.0f, Wc);
    q.sc = g.dv / b.cosa;
    if (GEOM == MBX_GEOM_CONE) {
        const float W = b.mag * g.dv_over_dr;
        const float z0 = g.slice_off - g.dv * g.slice_center;  // z of slice 0
        q.W = W;
        q.invW = 1.0f / W;
        q.a = (b.mag * z0 + g.row_off) / g.dr + g.row_center;
        q.hv = 0.5f * (W + 1.0f);
        q.Lv = fminf(1.0f, W);
        q.c0 = b.mag * z0 * g.inv_sdd;
        q.c1 = b.mag * g.dv * g.inv_sdd;
    } else {
        q.W = 1.0f;
        q.invW = 1.0f;
        q.a = 0.0f;
        q.hv = 1.0f;
        q.Lv = 1.0f;
        q.c0 = 0.0f;
        q.c1 = 0.0f;
    }
    return q;
}

// Horizontal-fan weight of channel n (ConeBeamModel.forward_horizontal_fan_pixel_batch_to_one_view,
// mbirjax/cone_beam.py:276-282): delta_voxel * clip((W_c+1)/2 - |n_p - n|, 0, min(1,W_c)) / cos(alpha),
// zero outside [0, C).
__device__ __forceinline__ float chan_weight(const PV &q, int n, int C) {
    const float L = clampf(q.hc - fabsf(q.np - (float)n), 0.0f, q.Lc);
    return (n >= 0 && n < C) ? q.sc * L : 0.0f;
}

// Vertical-fan weight between slice position t and detector row m
// (mbirjax/cone_beam.py:367-368, 529-531) without the 1/cos(phi) factor.
__device__ __forceinline__ float row_weight(const PV &q, float t_minus_m) {
    return clampf(q.hv - fabsf(t_minus_m), 0.0f, q.Lv);
}

}  // namespace mbx
