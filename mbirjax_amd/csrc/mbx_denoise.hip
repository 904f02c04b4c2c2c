// qGGMRF denoiser subset update and the 27-point median filter on gfx950.
//
// mbx_denoise_subset_update restates vcd_subset_denoiser
// (mbirjax/denoising.py:366-411): VCD with the identity as forward model, so
// the "sinogram" is the image and the forward Hessian is 1:
//   qGGMRF gradient/Hessian at the subset    (mbx_qggmrf_grad_hess)
//   delta = -((-e/sigma^2 + pg) / (1 + ph)), FP64 partials of
//     prior_linear = sum pg*delta, prior_quadratic = sum ph*delta^2 / 1.1,
//     forward_linear = sum e*delta / sigma^2, forward_quadratic = sum delta^2 / sigma^2
//   alpha = clip((fl - pl)/(fq + pq + eps32), eps32, 1.5)      (on the device)
//   x[S] += alpha*delta ; e[S] -= alpha*delta
// mbx_median_filter3d restates median_filter3d (mbirjax/denoising.py:414-496):
// the median of the 3x3x3 neighbourhood with edge-replicated borders.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "mbx_internal.hpp"

namespace mbx {

constexpr float kDnEps32 = 1.1920928955078125e-07f;  // jnp.finfo(jnp.float32).eps
constexpr int kDnBlocks = 1024;                      // partial-sum blocks
constexpr double kDnOverrelax = 1.1;                 // prior_overrelaxation_factor (denoising.py:386)

// delta and the FP64 partials [pl, pq (before 1/1.1), fl (before 1/sigma^2), fq (before 1/sigma^2)]
__global__ __launch_bounds__(256) void denoise_delta_kernel(const float *__restrict__ err,
                                                            const float *__restrict__ pg,
                                                            const float *__restrict__ ph,
                                                            const int32_t *__restrict__ pix, int npix, int Nz,
                                                            float fmc, float *__restrict__ delta,
                                                            double *__restrict__ part) {
    __shared__ double scratch[8];
    const int64_t n = (int64_t)npix * Nz;
    double pl = 0.0, pq = 0.0, fl = 0.0, fq = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / Nz), k = (int)(t - (int64_t)s * Nz);
        const float e = err[(size_t)pix[s] * Nz + k];
        const float fg = -(fmc * e);
        const float g = pg[t], h = ph[t];
        const float dl = -((fg + g) / (1.0f + h));
        delta[t] = dl;
        pl += (double)g * dl;
        pq += (double)h * dl * dl;
        fl += (double)e * dl;
        fq += (double)dl * dl;
    }
    const double r0 = block_reduce<double>(pl, scratch);
    const double r1 = block_reduce<double>(pq, scratch);
    const double r2 = block_reduce<double>(fl, scratch);
    const double r3 = block_reduce<double>(fq, scratch);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r0;
        part[kDnBlocks + blockIdx.x] = r1;
        part[2 * kDnBlocks + blockIdx.x] = r2;
        part[3 * kDnBlocks + blockIdx.x] = r3;
    }
}

// scal[0] = alpha; stats[1] += alpha, stats[2] += 1 (subsets done)
__global__ __launch_bounds__(256) void denoise_alpha_kernel(const double *__restrict__ part, int nblk, double fmc,
                                                            double *__restrict__ scal, double *__restrict__ stats) {
    __shared__ double scratch[8];
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a += part[i];
        b += part[kDnBlocks + i];
        c += part[2 * kDnBlocks + i];
        d += part[3 * kDnBlocks + i];
    }
    const double pl = block_reduce<double>(a, scratch);
    const double pq = block_reduce<double>(b, scratch) / kDnOverrelax;
    const double fl = fmc * block_reduce<double>(c, scratch);
    const double fq = fmc * block_reduce<double>(d, scratch);
    if (threadIdx.x == 0) {
        const double eps = (double)kDnEps32;
        double alpha = (fl - pl) / (fq + pq + eps);
        alpha = alpha < eps ? eps : (alpha > 1.5 ? 1.5 : alpha);
        if (alpha != alpha) alpha = eps;  // only if every term is NaN
        scal[0] = alpha;
        if (stats) {
            stats[1] += alpha;
            stats[2] += 1.0;
        }
    }
}

// x[S] += alpha*delta (a pixel listed twice gets both updates, as .at[].add);
// e[S] = e[S] - alpha*delta; per-block sum |alpha*delta|
__global__ __launch_bounds__(256) void denoise_update_kernel(float *__restrict__ x, float *__restrict__ err,
                                                             const int32_t *__restrict__ pix, int npix, int Nz,
                                                             const double *__restrict__ scal,
                                                             const float *__restrict__ delta,
                                                             double *__restrict__ part) {
    __shared__ double scratch[8];
    const float alpha = (float)scal[0];
    const int64_t n = (int64_t)npix * Nz;
    double l1 = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / Nz), k = (int)(t - (int64_t)s * Nz);
        const size_t o = (size_t)pix[s] * Nz + k;
        const float d = delta[t];
        const float ad = alpha * d;
        atomicAdd(&x[o], ad);
        // duplicates carry the same delta (same pixel, same error): one value is stored
        err[o] = err[o] - ad;
        l1 += fabs((double)ad);
    }
    const double r = block_reduce<double>(l1, scratch);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

struct DnLayout {
    size_t pg, ph, delta, part, scal, total;
};

static size_t dn_align(size_t x) { return (x + 255) & ~(size_t)255; }

static DnLayout dn_layout(int npix, int Nz) {
    DnLayout L;
    const size_t cyl = dn_align((size_t)npix * Nz * sizeof(float));
    L.pg = 0;
    L.ph = L.pg + cyl;
    L.delta = L.ph + cyl;
    L.part = L.delta + cyl;
    L.scal = L.part + dn_align(4 * (size_t)kDnBlocks * sizeof(double));
    L.total = L.scal + 256;
    return L;
}

static int dn_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(kDnBlocks, (n + 255) / 256)); }

// 27-point median with edge replication: the 14th smallest of the 27 values,
// by a partial selection sort on registers.
template <typename T>
__global__ __launch_bounds__(256) void median3_kernel(const T *__restrict__ x, int d0, int d1, int d2,
                                                      T *__restrict__ out) {
    const int64_t n = (int64_t)d0 * d1 * d2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int i2 = (int)(t % d2);
        const int64_t r = t / d2;
        const int i1 = (int)(r % d1), i0 = (int)(r / d1);
        T v[27];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int j0 = min(max(i0 + a - 1, 0), d0 - 1);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int j1 = min(max(i1 + b - 1, 0), d1 - 1);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int j2 = min(max(i2 + c - 1, 0), d2 - 1);
                    v[9 * a + 3 * b + c] = x[((size_t)j0 * d1 + j1) * d2 + j2];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 14; ++i)
#pragma unroll
            for (int j = i + 1; j < 27; ++j) {
                const T lo = v[i] < v[j] ? v[i] : v[j];
                const T hi = v[i] < v[j] ? v[j] : v[i];
                v[i] = lo;
                v[j] = hi;
            }
        out[t] = v[13];
    }
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int64_t mbx_denoise_workspace_bytes(int32_t npix, int32_t num_slices) {
    if (npix < 0 || num_slices <= 0) return -1;
    return (int64_t)dn_layout(npix, num_slices).total;
}

int mbx_denoise_subset_update(float *flat_image, float *flat_error, int32_t num_rows, int32_t num_cols,
                              int32_t num_slices, const int32_t *pix, int32_t npix, double fm_constant,
                              const mbx_qggmrf *qparams, double *stats, void *workspace, int64_t workspace_bytes,
                              void *stream) {
    if (num_rows <= 0 || num_cols <= 0 || num_slices <= 0 || npix <= 0)
        return fail(MBX_EINVAL, "image dimensions and npix must be positive");
    if (!flat_image || !flat_error || !pix) return fail(MBX_EINVAL, "NULL device pointer");
    if (!(fm_constant > 0)) return fail(MBX_EINVAL, "fm_constant must be positive");
    const DnLayout L = dn_layout(npix, num_slices);
    if (!workspace || workspace_bytes < (int64_t)L.total)
        return fail(MBX_EWORKSPACE, "denoiser workspace too small: need " + std::to_string(L.total) + " bytes");
    hipStream_t st = (hipStream_t)stream;
    char *ws = (char *)workspace;
    float *pg = (float *)(ws + L.pg), *ph = (float *)(ws + L.ph), *delta = (float *)(ws + L.delta);
    double *part = (double *)(ws + L.part), *scal = (double *)(ws + L.scal);
    int rc = mbx_qggmrf_grad_hess(flat_image, num_rows, num_cols, num_slices, pix, npix, qparams, pg, ph, stream);
    if (rc != MBX_OK) return rc;
    const int nb = dn_blocks((int64_t)npix * num_slices);
    hipLaunchKernelGGL(denoise_delta_kernel, dim3(nb), dim3(256), 0, st, flat_error, pg, ph, pix, npix, num_slices,
                       (float)fm_constant, delta, part);
    MBX_LAUNCH_CHECK();
    hipLaunchKernelGGL(denoise_alpha_kernel, dim3(1), dim3(256), 0, st, part, nb, fm_constant, scal, stats);
    MBX_LAUNCH_CHECK();
    hipLaunchKernelGGL(denoise_update_kernel, dim3(nb), dim3(256), 0, st, flat_image, flat_error, pix, npix, num_slices,
                       scal, delta, part);
    MBX_LAUNCH_CHECK();
    if (stats) return launch_sum_partials(part, nb, stats + 0, 1, st);
    return MBX_OK;
}

int mbx_median_filter3d(const void *x, int32_t dtype, int32_t d0, int32_t d1, int32_t d2, void *out, void *stream) {
    if (d0 <= 0 || d1 <= 0 || d2 <= 0) return fail(MBX_EINVAL, "median filter dimensions must be positive");
    if (!x || !out) return fail(MBX_EINVAL, "NULL device pointer");
    if (x == out) return fail(MBX_EINVAL, "median filter input and output must not alias");
    const int64_t n = (int64_t)d0 * d1 * d2;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == MBX_DTYPE_F32)
        hipLaunchKernelGGL((median3_kernel<float>), dim3(blocks), dim3(256), 0, st, (const float *)x, d0, d1, d2,
                           (float *)out);
    else if (dtype == MBX_DTYPE_F64)
        hipLaunchKernelGGL((median3_kernel<double>), dim3(blocks), dim3(256), 0, st, (const double *)x, d0, d1, d2,
                           (double *)out);
    else
        return fail(MBX_EINVAL, "median filter dtype must be MBX_DTYPE_F32 or MBX_DTYPE_F64");
    MBX_LAUNCH_CHECK();
    return MBX_OK;
}

}  // extern "C"
