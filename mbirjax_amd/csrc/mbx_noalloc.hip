// mbx_back_project_noalloc: the sparse back projection of mbx_back_project
// with its scratch in a file-scope device array instead of a caller
// workspace, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
// calls no hipMalloc, hipFree or allocating library routine.
//
// The scratch holds the geometry table of one pixel chunk and the back
// projection's view-group partials.  Neither needs clearing: the table kernel
// writes every record and n_c entry of the chunk, and every wave of the back
// projection writes its partials for every slice before the reduction reads
// them.  Each GPU holds its own copy of the array until the process ends;
// every call on a GPU uses that GPU's copy, so calls that overlap on one GPU
// (other streams, other host threads) race on it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "mbx_internal.hpp"

namespace mbx {

constexpr int64_t kScratchBytes = 64ll << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (each GPU holds the array at its own address).
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static int scratch_address(hipStream_t st, void **out) {
    int dev = 0;
    hipError_t e = st ? hipStreamGetDevice(st, &dev) : hipGetDevice(&dev);
    if (e != hipSuccess) return fail(MBX_ELAUNCH, std::string("no device for the stream: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= 64) return fail(MBX_EUNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        MBX_HIP_CHECK(hipGetDevice(&cur));
        if (cur != dev) MBX_HIP_CHECK(hipSetDevice(dev));
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(MBX_ELAUNCH, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MBX_OK;
}

// Pixels per table chunk whose table and partials fit the scratch: the chunk
// of mbx_back_project when it fits (the same plans, so bitwise the same
// sums), else the largest chunk whose table plus the partials of any pixel
// count up to it fit (both grow with the chunk: bisection).  0 = nothing fits.
static int scratch_chunk(const DevGeom &d, int npix, int nviews) {
    const int reg = table_chunk(d, npix, nviews);
    if (table_bytes(reg, nviews) + back_part_bytes(d, reg, nviews) <= kScratchBytes) return reg;
    auto fits = [&](int c) { return table_bytes(c, nviews) + back_part_bound(d, c, nviews) <= kScratchBytes; };
    if (!fits(1)) return 0;
    int lo = 1, hi = reg;
    while (lo < hi) {
        const int mid = lo + (hi - lo + 1) / 2;
        if (fits(mid)) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// out[i] = 0 for i < n: a kernel rather than a memset, so that the call can
// be captured into a HIP graph.
__global__ __launch_bounds__(256) void zero_fill_kernel(float *__restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = 0.0f;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int64_t mbx_noalloc_scratch_bytes(void) { return kScratchBytes; }

int mbx_back_project_noalloc(const mbx_geom *g, const float *angles, const void *sino, int32_t sino_dtype,
                             int32_t num_sino_views, const float *weights, const int32_t *pix, int32_t npix,
                             const int32_t *views, int32_t nviews, int32_t coeff_power, float scale,
                             int32_t accumulate, float *out, void *stream) {
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
    hipStream_t st = (hipStream_t)stream;
    if (npix == 0) return MBX_OK;
    if (nviews == 0) {
        if (!accumulate) {
            const int64_t n = (int64_t)npix * d.Nz;
            hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256),
                               0, st, out, n);
            MBX_LAUNCH_CHECK();
        }
        return MBX_OK;
    }
    const int chunk = scratch_chunk(d, npix, nviews);
    if (chunk == 0)
        return fail(MBX_EUNSUPPORTED, "mbx_back_project_noalloc: the geometry table of one pixel over " +
                                          std::to_string(nviews) + " views and its partials exceed the " +
                                          std::to_string(kScratchBytes >> 20) + " MiB scratch");
    void *base = nullptr;
    rc = scratch_address(st, &base);
    if (rc != MBX_OK) return rc;
    float *T = (float *)base;
    int *NC = table_nc(T, chunk, nviews);
    float *part = (float *)((char *)base + table_bytes(chunk, nviews));
    for (int c0 = 0; c0 < npix; c0 += chunk) {
        const int n = std::min(chunk, npix - c0);
        rc = launch_table(d, angles, pix + c0, n, views, nviews, T, NC, st);
        if (rc != MBX_OK) return rc;
        // every chunk planned as a full one, as mbx_back_project does
        rc = launch_back_tab(d, T, n, sino, sino_dtype, weights, views, nviews, coeff_power, scale, accumulate,
                             out + (size_t)c0 * d.Nz, part, st, nullptr, chunk);
        if (rc != MBX_OK) return rc;
    }
    return MBX_OK;
}

}  // extern "C"
