// Host-side helpers shared by the libmbx translation units: error reporting,
// argument validation and the launch functions that the C ABI wraps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

#include "../../include/mbx.h"
#include "mbx_device.hpp"

namespace mbx {

// Thread-local message behind mbx_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define MBX_HIP_CHECK(expr)                                                                         \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            return ::mbx::fail(MBX_ELAUNCH, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

#define MBX_LAUNCH_CHECK()                                                                              \
    do {                                                                                                \
        hipError_t _e = hipGetLastError();                                                              \
        if (_e != hipSuccess)                                                                           \
            return ::mbx::fail(MBX_ELAUNCH, std::string("kernel launch failed: ") + hipGetErrorString(_e)); \
    } while (0)

// Validates g and fills the float32 device copy.  Returns MBX_OK or an error.
int make_dev_geom(const mbx_geom *g, DevGeom *out);

inline int dtype_size(int dt) { return dt == MBX_DTYPE_BF16 || dt == MBX_DTYPE_F16 ? 2 : 4; }
// sinogram storage the back projector, transposes, dots and the error-sinogram update read
inline bool dtype_ok(int dt) { return dt == MBX_DTYPE_F32 || dt == MBX_DTYPE_BF16 || dt == MBX_DTYPE_F16; }
// storage the forward projector writes (and the delta sinogram of the VCD update)
inline bool dtype_fwd_ok(int dt) { return dt == MBX_DTYPE_F32 || dt == MBX_DTYPE_BF16; }

// Per-(view, pixel) geometry table shared by both projectors: one record of
// kRec floats per (view slot, pixel) (layout in mbx_project.hip).
constexpr int kRec = 16;
constexpr int kFill = 4;  // 64-element groups loaded per round (forward staging)
constexpr int kBackFill = 2;  // 64-row chunks loaded per round (back projection)
// MI355X (gfx950): 8 XCDs of 32 CUs, 160 KB of LDS per CU; consecutive
// workgroups of a launch go to consecutive XCDs (blockIdx % 8)
constexpr int kXcds = 8;
constexpr int kCusPerXcd = 32;
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr int64_t kTableBudget = 512ll << 20;
// Rows [first, last] packed in record word 15.
__device__ __forceinline__ int rec_mlo(float w) { return __float_as_int(w) & 0xffff; }
__device__ __forceinline__ int rec_mhi(float w) { return __float_as_int(w) >> 16; }

int table_chunk(const DevGeom &g, int npix, int nviews);  // pixels per table
int64_t table_bytes(int chunk, int nviews);                // records + n_c array
// Slot s of the table is pixel pix[s * pstride].
int launch_table(const DevGeom &g, const float *angles, const int32_t *pix, int npix, const int32_t *views, int nviews,
                 float *T, int *NC, hipStream_t st, int pstride = 1);
inline int *table_nc(float *T, int chunk, int nviews) {
    return (int *)((char *)T + (((int64_t)chunk * nviews * 16 * 4 + 255) & ~255ll));
}

// Back projection planning: number of view splits (partials) and slice chunking.
struct BackPlan {
    int nsplit;  // view groups reduced at the end (1 = write directly)
    int kpl;     // slices per lane (chunk = 64*kpl slices)
    int nchunk;  // slice chunks
    int cap;     // LDS floats per wave for the detector column
    int unit;    // every W = mag*dv/dr <= 2 with radius-1 row taps: the unit-sum weights (coeff_power 1)
};
BackPlan plan_back(const DevGeom &g, int npix, int nviews);
int64_t back_part_bytes(const DevGeom &g, int chunk, int nviews);  // view-group partials
// The most partials any pixel count up to n needs (the split grows as the
// count falls, so a smaller subset can need more than a larger one).
int64_t back_part_bound(const DevGeom &g, int n, int nviews);
int64_t back_workspace_bytes(const DevGeom &g, int npix, int nviews);

struct FwdPlan {
    int nw;       // waves per block (one slab of 64 * rpl rows each)
    int rpl;      // detector rows per lane (64 apart)
    int rows_pb;  // detector rows per block
    int nrb;      // row blocks
    int cb;       // channels per block
    int ncb;      // channel blocks
    int list;     // candidate list capacity
    int nb;       // candidates staged per batch
    int cstride;  // LDS floats per staged cylinder
    size_t lds;   // dynamic LDS bytes
    int64_t nblocks;
};
FwdPlan plan_fwd(const DevGeom &g, int nviews, int layout);
int64_t fwd_workspace_bytes(const DevGeom &g, int npix, int nviews);
// Forward scratch besides the geometry table (partials, binned pixel lists).
int64_t fwd_aux_bytes(const DevGeom &g, int chunk, int nviews);

// Launchers used by the C ABI and by the VCD subset update.  All arguments are
// already validated.  launch_back / launch_fwd build the geometry table in
// their workspace (chunking large pixel sets); the _tab variants take a
// table that is already built.
int launch_back(const DevGeom &g, const float *angles, const void *sino, int sino_dtype, const float *weights,
                const int32_t *pix, int npix, const int32_t *views, int nviews, int coeff_power,
                float scale, int accumulate, float *out, void *ws, hipStream_t st);
int launch_fwd(const DevGeom &g, const float *angles, const float *vox, const int32_t *pix, int npix,
               const int32_t *views, int nviews, void *sino, int sino_dtype, int layout, int accumulate,
               const float *weights, double *sumsq, void *ws, hipStream_t st);
// defer (may be NULL): when the views are split into groups, leave the nsplit
// partials (npix * Nz floats each, in part) unreduced and set *defer = nsplit
// (0 when out is written).  The reduction is the float sum over the groups in
// order, times scale.  plan_npix: plan the launch as if it had
// max(npix, plan_npix) pixels -- a chunk of a larger call passes the chunk
// size that back_part_bytes() sized part for (the split grows as the pixel
// count falls, so a smaller last chunk planned on its own could need more).
int launch_back_tab(const DevGeom &g, const float *T, int npix, const void *sino, int sino_dtype,
                    const float *weights, const int32_t *views, int nviews, int coeff_power, float scale,
                    int accumulate, float *out, float *part, hipStream_t st, int *defer = nullptr,
                    int plan_npix = 0);
// The cylinder of table slot s is vox + s * pstride * Nz.  steps: kFwdBin
// (bins and forward records from the table), kFwdProject (the projection
// from bins and records already in aux), or both; kFwdPartials: the
// fwd_partials() FP64 partials (one per wave) of the weighted squared norm
// are written to the start of aux (and not reduced unless sumsq is given).
constexpr int kFwdBin = 1, kFwdProject = 2, kFwdPartials = 4;
int64_t fwd_partials(const DevGeom &g, int nviews, int layout);
int launch_fwd_tab(const DevGeom &g, const float *T, const int *NC, const float *vox, int npix, int nviews, void *sino,
                   int sino_dtype, int layout, int accumulate, const float *weights, void *aux, int chunk,
                   double *sumsq, hipStream_t st, int pstride = 1, int steps = kFwdBin | kFwdProject);
// Sum `n` doubles in fixed order into out[0] (one block).
int launch_sum_partials(const double *partials, int64_t n, double *out, int accumulate, hipStream_t st);

}  // namespace mbx
