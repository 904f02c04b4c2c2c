// qGGMRF prior, FP64 reductions and the VCD subset update on gfx950.
//
// mbx_vcd_subset_update restates vcd_subset_updater
// (mbirjax/tomography_model.py:1343-1542) as a fixed kernel chain on one
// stream with no host synchronisation:
//   geometry table of the subset
//   back projection of W*e            -> forward gradient (before -1/sigma_y^2),
//                                        as view-group partials
//   [forward binning, on a side stream meanwhile for large subsets]
//   delta kernel: sum of the partials, qGGMRF (or prox) gradient/Hessian,
//   delta = -(fg + pg)/(fh + ph)      -> with FP64 partials of prior_linear,
//                                        prior_quadratic and forward_linear
//   forward projection of delta       -> delta sinogram, FP64 partials of
//                                        forward_quadratic
//   alpha = clip((fl - pl)/(fq + pq + eps32), eps32, 1.5)   (on the device)
//   [positivity clamp + re-projection]
//   x[S] += alpha*delta ; e -= alpha*delta_sino
// forward_linear uses the adjoint identity
//   sum(W e . A delta) = <A^T (W e), delta>,
// so it is a by-product of the back projection instead of an extra pass over
// the sinogram (get_forward_lin_quad, mbirjax/tomography_model.py:1546-1631).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "mbx_internal.hpp"

namespace mbx {

// A second stream per device (default priority) that runs the forward binning
// of a subset while the back projection runs on the caller's stream, with the
// two events of the fork and the join.  Created on first use; the mutex
// serialises the enqueueing of concurrent calls, which share the events.
struct SideStream {
    std::mutex mu;
    bool tried = false;
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
static SideStream g_side[64];

// The side stream of the device that runs st (the current device for the
// null stream), or NULL when there is none.
static SideStream *side_stream(hipStream_t st) {
    int dev = 0;
    const hipError_t e = st ? hipStreamGetDevice(st, &dev) : hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return nullptr;
    }
    SideStream &ss = g_side[dev];
    std::lock_guard<std::mutex> lk(ss.mu);
    if (!ss.tried) {
        ss.tried = true;
        int cur = dev;
        bool ok = hipGetDevice(&cur) == hipSuccess && (cur == dev || hipSetDevice(dev) == hipSuccess);
        ok = ok && hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&ss.join, hipEventDisableTiming) == hipSuccess;
        if (cur != dev) (void)hipSetDevice(cur);
        if (!ok) {
            (void)hipGetLastError();
            ss.s = nullptr;  // then the binning stays on the caller's stream
        }
    }
    return ss.s ? &ss : nullptr;
}

constexpr float kEps32 = 1.1920928955078125e-07f;  // jnp.finfo(jnp.float32).eps
constexpr int kRedBlocks = 1024;

// ======================================================================
// qGGMRF gradient and Hessian (mbirjax/qggmrf.py:118-279)
// ======================================================================

struct QP {
    float b[6];
    float tsig;      // T*sigma_x
    float inv_tsig;  // 1 / (T*sigma_x)
    float scale;     // T^(p-2) / sigma_x^2
    float qm2;       // q - 2
    float qmp;       // q - p
    float q_over_p;  // q / p
};

// get_2_b_tilde (mbirjax/qggmrf.py:250-279): b * rho'(delta)/delta.
// sd >= eps32 is a normal number, so the hardware log2 needs no denormal
// scaling, and sd^(q-2), sd^(q-p) (1 <= q <= p <= 2) are normal unless
// |delta| / (T sigma_x) exceeds 2^126, so neither does the hardware exp2; the
// quotients use the hardware reciprocal (1 ulp).
__device__ __forceinline__ float two_b_tilde(float delta, float bval, const QP &q) {
    const float sd = fabsf(delta) * q.inv_tsig + kEps32;
    const float l2 = __builtin_amdgcn_logf(sd);
    const float d_q2 = __builtin_amdgcn_exp2f(q.qm2 * l2);  // sd^(q-2)
    const float d_qp = __builtin_amdgcn_exp2f(q.qmp * l2);  // sd^(q-p)
    const float num = d_q2 * (q.q_over_p + d_qp);
    const float den1 = 1.0f + d_qp;
    const float den = den1 * den1;
    return bval * (q.scale * num * __builtin_amdgcn_rcpf(den));
}

// Surrogate gradient and Hessian of voxel (pixel p, slice k) over its six
// neighbours.
__device__ __forceinline__ void qggmrf_gh(const float *__restrict__ x, int rows, int cols, int Nz, int p, int k,
                                          const QP &q, float &g, float &h) {
    const int i = p / cols, j = p - i * cols;
    const float *xc = x + (size_t)p * Nz;
    const float v = xc[k];
    // cylinder neighbours with reflected (zero-difference) ends (qggmrf.py:151-188)
    const float dplus = (k + 1 < Nz) ? xc[k + 1] - v : 0.0f;
    const float dminus = (k > 0) ? v - xc[k - 1] : 0.0f;
    const float tp = two_b_tilde(dplus, 1.0f, q);
    const float tm = two_b_tilde(dminus, 1.0f, q);
    g = -q.b[4] * (tp * dplus);
    h = q.b[4] * tp;
    g += q.b[5] * (tm * dminus);
    h += q.b[5] * tm;
    // in-plane neighbours [row+1, row-1, col+1, col-1], clipped at the borders
    // (ravel_multi_index(..., mode='clip'), qggmrf.py:191-247)
    const int nb[4] = {min(i + 1, rows - 1) * cols + j, max(i - 1, 0) * cols + j, i * cols + min(j + 1, cols - 1),
                       i * cols + max(j - 1, 0)};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float d = v - x[(size_t)nb[t] * Nz + k];
        const float bt = two_b_tilde(d, q.b[t], q);
        g += bt * d;
        h += bt;
    }
}

__global__ __launch_bounds__(256) void qggmrf_kernel(const float *__restrict__ x, int rows, int cols, int Nz,
                                                     const int32_t *__restrict__ pix, int npix, QP q,
                                                     float *__restrict__ grad, float *__restrict__ hess) {
    const int s = blockIdx.x;
    if (s >= npix) return;
    const int p = pix[s];
    for (int k = threadIdx.x; k < Nz; k += blockDim.x) {
        float g, h;
        qggmrf_gh(x, rows, cols, Nz, p, k, q, g, h);
        grad[(size_t)s * Nz + k] = g;
        hess[(size_t)s * Nz + k] = h;
    }
}

static QP make_qp(const mbx_qggmrf *P) {
    QP q;
    for (int t = 0; t < 6; ++t) q.b[t] = (float)P->b[t];
    q.tsig = (float)(P->T * P->sigma_x);
    q.inv_tsig = (float)(1.0 / (P->T * P->sigma_x));
    q.scale = (float)(std::pow(P->T, P->p - 2.0) / (P->sigma_x * P->sigma_x));
    q.qm2 = (float)(P->q - 2.0);
    q.qmp = (float)(P->q - P->p);
    q.q_over_p = (float)(P->q / P->p);
    return q;
}

static int check_qparams(const mbx_qggmrf *P) {
    if (!P) return fail(MBX_EINVAL, "qggmrf params pointer is NULL");
    if (!(P->sigma_x > 0) || !(P->T > 0) || !(P->p > 0) || !(P->q > 0))
        return fail(MBX_EINVAL, "qggmrf sigma_x, T, p and q must be positive");
    return MBX_OK;
}

// prox_gradient_at_indices (mbirjax/qggmrf.py:8-28)
__global__ __launch_bounds__(256) void prox_grad_kernel(const float *__restrict__ x, const float *__restrict__ prox,
                                                        const int32_t *__restrict__ pix, int npix, int Nz,
                                                        float inv_sp2, float *__restrict__ grad) {
    const int64_t n = (int64_t)npix * Nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / Nz), k = (int)(t - (int64_t)s * Nz);
        const size_t o = (size_t)pix[s] * Nz + k;
        grad[t] = inv_sp2 * (x[o] - prox[o]);
    }
}

// ======================================================================
// Reductions
// ======================================================================


__global__ __launch_bounds__(256) void sum_partials_kernel(const double *__restrict__ part, int64_t n,
                                                           double *__restrict__ out, int accumulate) {
    __shared__ double scratch[8];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    const double r = block_reduce<double>(s, scratch);
    if (threadIdx.x == 0) out[0] = accumulate ? out[0] + r : r;
}

int launch_sum_partials(const double *partials, int64_t n, double *out, int accumulate, hipStream_t st) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partials, n, out, accumulate);
    MBX_LAUNCH_CHECK();
    return MBX_OK;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void dot_kernel(const TA *__restrict__ a, const TB *__restrict__ b,
                                                  const float *__restrict__ w, int64_t n, double *__restrict__ part) {
    __shared__ double scratch[8];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double av = (double)to_float(a[i]);
        const double bv = b ? (double)to_float(b[i]) : av;
        s += w ? av * bv * (double)w[i] : av * bv;
    }
    const double r = block_reduce<double>(s, scratch);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <typename TA>
static void launch_dot_a(const TA *a, const void *b, int bdt, const float *w, int64_t n, int nb, double *part,
                         hipStream_t st) {
    if (bdt == MBX_DTYPE_F32)
        hipLaunchKernelGGL((dot_kernel<TA, float>), dim3(nb), dim3(256), 0, st, a, (const float *)b, w, n, part);
    else if (bdt == MBX_DTYPE_BF16)
        hipLaunchKernelGGL((dot_kernel<TA, bf16>), dim3(nb), dim3(256), 0, st, a, (const bf16 *)b, w, n, part);
    else
        hipLaunchKernelGGL((dot_kernel<TA, f16>), dim3(nb), dim3(256), 0, st, a, (const f16 *)b, w, n, part);
}

// e = e - float(alpha) * d: thread t updates elements 4t .. 4t+3 with streaming
// (nontemporal) accesses, the grid covering n/4 threads; the remaining (or, for
// unaligned arrays, all) elements go element-wise over the grid.
template <typename TE, typename TD>
__device__ __forceinline__ void axpy_body(TE *__restrict__ e, const TD *__restrict__ d, float a, int64_t n) {
    // vector path only when both arrays are aligned for the 4-element access
    const bool vec = ((uintptr_t)e % (4 * sizeof(TE)) == 0) && ((uintptr_t)d % (4 * sizeof(TD)) == 0);
    const int64_t n4 = vec ? (n >> 2) : 0;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n4) {
        float ev[4], dv[4];
        load4_nt(e + 4 * t, ev);
        load4_nt(d + 4 * t, dv);
#pragma unroll
        for (int u = 0; u < 4; ++u) ev[u] = fmaf(-a, dv[u], ev[u]);
        store4_nt(e + 4 * t, ev);
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = 4 * n4 + t; i < n; i += stride) e[i] = from_float<TE>(fmaf(-a, to_float(d[i]), to_float(e[i])));
}

// one thread per four elements
static int64_t axpy_blocks(int64_t n) { return (n / 4 + 255) / 256 + 1; }

template <typename TE, typename TD>
__global__ __launch_bounds__(256) void axpy_kernel(TE *__restrict__ e, const TD *__restrict__ d,
                                                   const double *__restrict__ alpha, int64_t n) {
    axpy_body(e, d, (float)alpha[0], n);
}

template <typename TE>
static void launch_axpy_e(TE *e, const void *d, int ddt, const double *alpha, int64_t n, hipStream_t st) {
    const int64_t blocks = axpy_blocks(n);
    if (ddt == MBX_DTYPE_F32)
        hipLaunchKernelGGL((axpy_kernel<TE, float>), dim3((unsigned)blocks), dim3(256), 0, st, e, (const float *)d,
                           alpha, n);
    else
        hipLaunchKernelGGL((axpy_kernel<TE, bf16>), dim3((unsigned)blocks), dim3(256), 0, st, e, (const bf16 *)d,
                           alpha, n);
}

// e (edt: f32, bf16 or f16) -= float(alpha[0]) * d (ddt: f32 or bf16)
static int launch_axpy(void *e, int edt, const void *d, int ddt, const double *alpha, int64_t n, hipStream_t st) {
    if (n == 0) return MBX_OK;
    if (edt == MBX_DTYPE_F32)
        launch_axpy_e((float *)e, d, ddt, alpha, n, st);
    else if (edt == MBX_DTYPE_BF16)
        launch_axpy_e((bf16 *)e, d, ddt, alpha, n, st);
    else
        launch_axpy_e((f16 *)e, d, ddt, alpha, n, st);
    MBX_LAUNCH_CHECK();
    return MBX_OK;
}

// ======================================================================
// VCD subset update pieces
// ======================================================================

// delta = -((fg + pg) / (fh + ph)) with fg = -fmc*bp, fh = fmc*H[p];
// partials: [0] prior_linear = sum pg*delta, [1] sum ph*delta^2, [2] sum bp*delta
// QG: the prior gradient and Hessian are the qGGMRF surrogate's, computed here
// from the recon x (else pg is read and ph = ph_const, the proximal map's).
// bsplit > 0: bp is the float sum, in order, of the back projection's bsplit
// view-group partials in bpart (each npix * Nz).
template <bool QG>
__global__ __launch_bounds__(256) void delta_kernel(const float *__restrict__ bp, const float *__restrict__ bpart,
                                                    int bsplit, const float *__restrict__ pg,
                                                    float ph_const, const float *__restrict__ x, int rows, int cols, QP q,
                                                    const float *__restrict__ fm_hess, const int32_t *__restrict__ pix,
                                                    int npix, int Nz, float fmc, float *__restrict__ delta,
                                                    double *__restrict__ part) {
    __shared__ double scratch[8];
    const int64_t n = (int64_t)npix * Nz;
    double pl = 0.0, pq = 0.0, fl = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / Nz), k = (int)(t - (int64_t)s * Nz);
        float b;
        if (bsplit > 0) {
            b = 0.0f;
#pragma unroll 8
            for (int j = 0; j < bsplit; ++j) b += bpart[(size_t)j * n + t];  // loads in flight, adds in order
        } else {
            b = bp[t];
        }
        const float fg = -(fmc * b);
        const float fh = fmc * fm_hess[(size_t)pix[s] * Nz + k];
        float g, h;
        if (QG) {
            qggmrf_gh(x, rows, cols, Nz, pix[s], k, q, g, h);
        } else {
            h = ph_const;
            g = pg[t];
        }
        const float dl = -((fg + g) / (fh + h));
        delta[t] = dl;
        pl += (double)g * dl;
        pq += (double)h * dl * dl;
        fl += (double)b * dl;
    }
    const double r0 = block_reduce<double>(pl, scratch);
    const double r1 = block_reduce<double>(pq, scratch);
    const double r2 = block_reduce<double>(fl, scratch);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r0;
        part[kRedBlocks + blockIdx.x] = r1;
        part[2 * kRedBlocks + blockIdx.x] = r2;
    }
}

// fq_raw = sum of delta_sino^2 * w, from the forward projection's nfq
// per-block partials fqpart (kept in scal[0]).
// alpha = clip((fl - pl) / (fq + pq + eps32), eps32, 1.5)   (tomography_model.py:1467-1471)
__global__ __launch_bounds__(1024) void alpha_kernel(const double *__restrict__ part, int nblk,
                                                    const double *__restrict__ fqpart, int64_t nfq, double fmc,
                                                    double *__restrict__ scal, double *__restrict__ stats) {
    __shared__ double scratch[16][4];  // one row per wave of the 1024-thread block
    double a = 0.0, b = 0.0, c = 0.0, f = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a += part[i];
        b += part[kRedBlocks + i];
        c += part[2 * kRedBlocks + i];
    }
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < nfq; i += blockDim.x) f += fqpart[i];
    // the four sums reduced together (each in the order block_reduce uses:
    // a shuffle tree per wave, then the waves in order)
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
        f += __shfl_down(f, off, 64);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        scratch[wave][0] = a;
        scratch[wave][1] = b;
        scratch[wave][2] = c;
        scratch[wave][3] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0, sc = 0.0, sf = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            sa += scratch[i][0];
            sb += scratch[i][1];
            sc += scratch[i][2];
            sf += scratch[i][3];
        }
        const double pl = sa, pq = 0.5 * sb, fl = fmc * sc, fq_raw = sf;
        scal[0] = fq_raw;
        const double fq = fmc * fq_raw;
        const double eps = (double)kEps32;
        double alpha = (fl - pl) / (fq + pq + eps);
        alpha = alpha < eps ? eps : (alpha > 1.5 ? 1.5 : alpha);
        if (alpha != alpha) alpha = eps;  // only if every term is NaN
        scal[1] = alpha;
        scal[2] = fl;
        scal[3] = fq;
        scal[4] = pl;
        scal[5] = pq;
        if (stats) {
            stats[1] += alpha;
            stats[2] += 1.0;
        }
    }
}

// delta = max(-x[S]/(alpha + eps32), delta)   (tomography_model.py:1497-1504)
__global__ __launch_bounds__(256) void positivity_kernel(const float *__restrict__ x, const int32_t *__restrict__ pix,
                                                         int npix, int Nz, const double *__restrict__ scal,
                                                         float *__restrict__ delta) {
    const float alpha = (float)scal[1];
    const float pc = 1.0f / (alpha + kEps32);
    const int64_t n = (int64_t)npix * Nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / Nz), k = (int)(t - (int64_t)s * Nz);
        delta[t] = fmaxf(-pc * x[(size_t)pix[s] * Nz + k], delta[t]);
    }
}

// x[S] += alpha*delta ; per-block sum |alpha*delta|   (tomography_model.py:1512-1536, 1820-1823)
__global__ __launch_bounds__(256) void update_recon_kernel(float *__restrict__ x, const int32_t *__restrict__ pix,
                                                           int npix, int Nz, const double *__restrict__ scal,
                                                           const float *__restrict__ delta, double *__restrict__ part) {
    __shared__ double scratch[8];
    const float alpha = (float)scal[1];
    const int64_t n = (int64_t)npix * Nz;
    double l1 = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t / Nz), k = (int)(t - (int64_t)s * Nz);
        const float ad = alpha * delta[t];
        // a subset can list a pixel twice when the padding outnumbers a subset
        // (vcd_utils.py:212-217); .at[].add then adds both copies
        atomicAdd(&x[(size_t)pix[s] * Nz + k], ad);
        l1 += fabs((double)ad);
    }
    const double r = block_reduce<double>(l1, scratch);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}


struct VcdLayout {
    int chunk;  // pixels per geometry table
    size_t table, bpart, fpart, bp, pg, delta, dsino, dpart, l1part, scal, total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static VcdLayout vcd_layout(const DevGeom &g, int npix, int ddt) {
    VcdLayout L;
    L.chunk = table_chunk(g, std::max(npix, 1), g.V);

    size_t off = 0;
    const size_t cyl = (size_t)npix * g.Nz * sizeof(float);
    L.table = off;
    off += align256((size_t)table_bytes(L.chunk, g.V));
    L.bpart = off;  // sized for every subset of up to npix pixels, so the whole
                    // layout grows with npix: a workspace sized for the largest
                    // subset of a set of partitions holds every smaller one
    off += align256((size_t)back_part_bound(g, L.chunk, g.V));
    L.fpart = off;
    off += align256((size_t)fwd_aux_bytes(g, L.chunk, g.V));
    L.bp = off;
    off += align256(cyl);
    L.pg = off;  // proximal-map prior gradient
    off += align256(cyl);
    L.delta = off;
    off += align256(cyl);
    L.dsino = off;
    off += align256((size_t)g.V * g.C * g.R * dtype_size(ddt));
    L.dpart = off;
    off += align256(3 * kRedBlocks * sizeof(double));
    L.l1part = off;
    off += align256(kRedBlocks * sizeof(double));
    L.scal = off;
    off += align256(16 * sizeof(double));
    L.total = off;
    return L;
}

static int red_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, kRedBlocks)); }

}  // namespace mbx

using namespace mbx;

extern "C" {

int mbx_qggmrf_grad_hess(const float *flat_recon, int32_t num_rows, int32_t num_cols, int32_t num_slices,
                         const int32_t *pix, int32_t npix, const mbx_qggmrf *params, float *grad, float *hess,
                         void *stream) {
    if (num_rows <= 0 || num_cols <= 0 || num_slices <= 0 || npix < 0)
        return fail(MBX_EINVAL, "bad recon shape or npix");
    int rc = check_qparams(params);
    if (rc != MBX_OK) return rc;
    if (npix == 0) return MBX_OK;
    if (!flat_recon || !pix || !grad || !hess) return fail(MBX_EINVAL, "NULL device pointer");
    const QP q = make_qp(params);
    const int threads = num_slices >= 256 ? 256 : (num_slices > 64 ? 128 : 64);
    hipLaunchKernelGGL(qggmrf_kernel, dim3(npix), dim3(threads), 0, (hipStream_t)stream, flat_recon, num_rows,
                       num_cols, num_slices, pix, npix, q, grad, hess);
    MBX_LAUNCH_CHECK();
    return MBX_OK;
}

int64_t mbx_dot_workspace_bytes(void) { return (int64_t)kRedBlocks * sizeof(double); }

int mbx_dot(const void *a, int32_t a_dtype, const void *b, int32_t b_dtype, const float *w, int64_t n, double *out,
            void *workspace, int64_t workspace_bytes, void *stream) {
    if (n < 0) return fail(MBX_EINVAL, "n must be >= 0");
    if (!dtype_ok(a_dtype) || (b && !dtype_ok(b_dtype))) return fail(MBX_EINVAL, "unknown dtype");
    if (!out || (n > 0 && !a)) return fail(MBX_EINVAL, "NULL device pointer");
    if (!workspace || workspace_bytes < mbx_dot_workspace_bytes())
        return fail(MBX_EWORKSPACE, "mbx_dot workspace too small");
    hipStream_t st = (hipStream_t)stream;
    const int nb = red_blocks(n);
    double *part = (double *)workspace;
    if (!b) b_dtype = a_dtype;
    if (a_dtype == MBX_DTYPE_F32)
        launch_dot_a((const float *)a, b, b_dtype, w, n, nb, part, st);
    else if (a_dtype == MBX_DTYPE_BF16)
        launch_dot_a((const bf16 *)a, b, b_dtype, w, n, nb, part, st);
    else
        launch_dot_a((const f16 *)a, b, b_dtype, w, n, nb, part, st);
    MBX_LAUNCH_CHECK();
    return launch_sum_partials(part, nb, out, 0, st);
}

int mbx_sino_axpy(void *e, int32_t e_dtype, const void *d, int32_t d_dtype, const double *alpha, int64_t n,
                  void *stream) {
    if (n < 0) return fail(MBX_EINVAL, "n must be >= 0");
    if (!dtype_ok(e_dtype) || !dtype_fwd_ok(d_dtype)) return fail(MBX_EINVAL, "unknown dtype");
    if (n > 0 && (!e || !d || !alpha)) return fail(MBX_EINVAL, "NULL device pointer");
    return launch_axpy(e, e_dtype, d, d_dtype, alpha, n, (hipStream_t)stream);
}

int64_t mbx_vcd_workspace_bytes(const mbx_geom *g, int32_t npix, int32_t delta_dtype) {
    DevGeom d;
    if (make_dev_geom(g, &d) != MBX_OK || npix < 0 || !dtype_fwd_ok(delta_dtype)) return -1;
    return (int64_t)vcd_layout(d, npix, delta_dtype).total;
}

}  // extern "C"

namespace mbx {

// The subset update; plan (nullable): the subset's forward plan over every
// view (mbx_forward_plan), whose table, bins and records it uses instead of
// building them.
static int vcd_update(const mbx_geom *g, const float *angles, float *flat_recon, void *e, int32_t e_dtype,
                      const float *weights, const float *fm_hessian, const int32_t *pix, int32_t npix,
                      double fm_constant, const mbx_qggmrf *qparams, const float *prox_input, double sigma_prox,
                      int32_t positivity, int32_t delta_dtype, void *plan, int64_t plan_bytes, double *stats,
                      void *workspace, int64_t workspace_bytes, void *stream) {
    DevGeom d;
    int rc = make_dev_geom(g, &d);
    if (rc != MBX_OK) return rc;
    if (npix <= 0) return fail(MBX_EINVAL, "npix must be positive");
    if (!dtype_ok(e_dtype) || !dtype_fwd_ok(delta_dtype)) return fail(MBX_EINVAL, "unknown dtype");
    if (!angles || !flat_recon || !e || !fm_hessian || !pix) return fail(MBX_EINVAL, "NULL device pointer");
    if (!prox_input) {
        rc = check_qparams(qparams);
        if (rc != MBX_OK) return rc;
    } else if (!(sigma_prox > 0)) {
        return fail(MBX_EINVAL, "sigma_prox must be positive");
    }
    if (!(fm_constant > 0)) return fail(MBX_EINVAL, "fm_constant must be positive");
    const VcdLayout L = vcd_layout(d, npix, delta_dtype);
    if (!workspace || workspace_bytes < (int64_t)L.total)
        return fail(MBX_EWORKSPACE, "VCD workspace too small: need " + std::to_string(L.total) + " bytes");
    hipStream_t st = (hipStream_t)stream;
    char *ws = (char *)workspace;
    float *bp = (float *)(ws + L.bp);
    float *pg = (float *)(ws + L.pg);
    float *delta = (float *)(ws + L.delta);
    void *dsino = (void *)(ws + L.dsino);
    double *dpart = (double *)(ws + L.dpart);
    double *l1part = (double *)(ws + L.l1part);
    double *scal = (double *)(ws + L.scal);
    const int64_t ncyl = (int64_t)npix * d.Nz;
    const int64_t nsino = (int64_t)d.V * d.C * d.R;
    const int nb = red_blocks(ncyl);

    float *T = (float *)(ws + L.table);
    float *bpart = (float *)(ws + L.bpart);
    void *fpart = ws + L.fpart;
    const bool one_table = npix <= L.chunk;  // build the table once for all projections
    if (plan) {
        // the plan's layout is the one mbx_forward_plan writes for these npix
        // pixels and all views: the table (records, then n_c) and the forward
        // aux (FP64 partial slots, pixel list, bucket starts, records)
        if (!one_table)
            return fail(MBX_EUNSUPPORTED, "a planned subset update needs a subset of one geometry-table chunk (" +
                                              std::to_string(L.chunk) + " pixels)");
        const int64_t need = fwd_workspace_bytes(d, npix, d.V);
        if (plan_bytes != need)
            return fail(MBX_EINVAL, "plan_bytes " + std::to_string(plan_bytes) + " is not the size " +
                                        std::to_string(need) + " of a forward plan of " + std::to_string(npix) +
                                        " pixels over all views");
        T = (float *)plan;
        fpart = (char *)plan + table_bytes(L.chunk, d.V);
    }
    int *NC = table_nc(T, L.chunk, d.V);

    // (c) back projection of the weighted error sinogram (with one table, the
    // delta kernel sums its view-group partials, and the forward binning runs
    // on the side stream meanwhile when it is long enough to pay for the fork
    // and join: 2^18 (view, pixel) records or more)
    const bool side_bin = !plan && one_table && (int64_t)npix * d.V >= (1 << 18);
    int bsplit = 0;
    SideStream *side = nullptr;
    std::unique_lock<std::mutex> side_lock;
    // the caller's stream waits for the side stream's work on every return
    // path once it is forked, so nothing of this call runs past its end: through
    // the join event once it is recorded, else (an error between the fork and
    // the record) by draining the side stream before returning
    struct Join {
        SideStream *&side;
        hipStream_t st;
        bool forked = false, recorded = false;
        ~Join() {
            if (forked && recorded) (void)hipStreamWaitEvent(st, side->join, 0);
            else if (forked) (void)hipStreamSynchronize(side->s);
        }
    } join{side, st};
    for (int c0 = 0; c0 < npix; c0 += L.chunk) {
        const int n = std::min(L.chunk, npix - c0);
        if (!plan) {
            rc = launch_table(d, angles, pix + c0, n, nullptr, d.V, T, NC, st);
            if (rc != MBX_OK) return rc;
        }
        if (side_bin && (side = side_stream(st)) != nullptr) {
            side_lock = std::unique_lock<std::mutex>(side->mu);
            MBX_HIP_CHECK(hipEventRecord(side->fork, st));
            MBX_HIP_CHECK(hipStreamWaitEvent(side->s, side->fork, 0));
            join.forked = true;  // from here on the side stream may hold work of this call
            rc = launch_fwd_tab(d, T, NC, delta, n, d.V, dsino, delta_dtype, MBX_LAYOUT_VCR, 0, weights, fpart,
                                L.chunk, nullptr, side->s, 1, kFwdBin);
            if (rc != MBX_OK) return rc;
            MBX_HIP_CHECK(hipEventRecord(side->join, side->s));
            join.recorded = true;
        }
        rc = launch_back_tab(d, T, n, e, e_dtype, weights, nullptr, d.V, 1, 1.0f, 0, bp + (size_t)c0 * d.Nz,
                             bpart, st, one_table ? &bsplit : nullptr, L.chunk);
        if (rc != MBX_OK) return rc;
    }
    // forward projection of delta over the subset (into dsino), optionally with forward_quadratic
    // (interleaved chunks: chunk c holds pixels c, c + nch, ..., so that every
    // chunk spreads over the image and the forward's channel blocks stay balanced)
    // (a second projection of one table chunk reuses the bins and records of the
    // first; with partials, the last chunk leaves the FP64 partials of
    // sum w * (A delta)^2 at the start of fpart)
    bool binned = plan != nullptr;
    auto project_delta = [&](bool partials, const float *w) -> int {
        const int nch = (npix + L.chunk - 1) / L.chunk;
        for (int c = 0; c < nch; ++c) {
            const int n = (npix - c + nch - 1) / nch;
            if (!one_table) {
                int r = launch_table(d, angles, pix + c, n, nullptr, d.V, T, NC, st, nch);
                if (r != MBX_OK) return r;
            }
            int r = launch_fwd_tab(d, T, NC, delta + (size_t)c * d.Nz, n, d.V, dsino, delta_dtype, MBX_LAYOUT_VCR,
                                   c > 0, w, fpart, L.chunk, nullptr, st, nch,
                                   (binned ? kFwdProject : kFwdBin | kFwdProject) |
                                       (partials && c == nch - 1 ? kFwdPartials : 0));
            if (r != MBX_OK) return r;
        }
        binned = one_table;
        return MBX_OK;
    };
    // (a) prior gradient and Hessian (the qGGMRF surrogate's inside the delta
    // kernel) and (d,e,f) delta and FP64 partials
    if (prox_input) {
        hipLaunchKernelGGL(prox_grad_kernel, dim3(nb), dim3(256), 0, st, flat_recon, prox_input, pix, npix, d.Nz,
                           (float)(1.0 / (sigma_prox * sigma_prox)), pg);
        MBX_LAUNCH_CHECK();
        hipLaunchKernelGGL((delta_kernel<false>), dim3(nb), dim3(256), 0, st, bp, bpart, bsplit, pg,
                           (float)(sigma_prox * sigma_prox), nullptr, d.rows, d.cols, QP{}, fm_hessian, pix, npix,
                           d.Nz, (float)fm_constant, delta, dpart);
    } else {
        hipLaunchKernelGGL((delta_kernel<true>), dim3(nb), dim3(256), 0, st, bp, bpart, bsplit, nullptr, 0.0f,
                           flat_recon, d.rows, d.cols, make_qp(qparams), fm_hessian, pix, npix, d.Nz,
                           (float)fm_constant, delta, dpart);
    }
    MBX_LAUNCH_CHECK();
    if (join.forked) {  // join: the bins and records are in fpart
        join.forked = join.recorded = false;
        MBX_HIP_CHECK(hipStreamWaitEvent(st, side->join, 0));
        side_lock.unlock();
        binned = true;
    }
    // (g,h) delta sinogram and forward_quadratic
    rc = project_delta(true, weights);
    if (rc != MBX_OK) return rc;
    // (i) step size on the device
    hipLaunchKernelGGL(alpha_kernel, dim3(1), dim3(1024), 0, st, dpart, nb, (const double *)fpart,
                       fwd_partials(d, d.V, MBX_LAYOUT_VCR), fm_constant, scal, stats);
    MBX_LAUNCH_CHECK();
    // (j) positivity
    if (positivity) {
        hipLaunchKernelGGL(positivity_kernel, dim3(nb), dim3(256), 0, st, flat_recon, pix, npix, d.Nz, scal, delta);
        MBX_LAUNCH_CHECK();
        rc = project_delta(false, nullptr);
        if (rc != MBX_OK) return rc;
    }
    // (k,m) recon update and ell1
    hipLaunchKernelGGL(update_recon_kernel, dim3(nb), dim3(256), 0, st, flat_recon, pix, npix, d.Nz, scal, delta,
                       l1part);
    MBX_LAUNCH_CHECK();
    if (stats) {
        rc = launch_sum_partials(l1part, nb, stats + 0, 1, st);
        if (rc != MBX_OK) return rc;
    }
    // (l) error sinogram update, alpha read from scal[1]
    return launch_axpy(e, e_dtype, dsino, delta_dtype, scal + 1, nsino, st);
}

}  // namespace mbx

extern "C" {

int mbx_vcd_subset_update(const mbx_geom *g, const float *angles, float *flat_recon, void *e, int32_t e_dtype,
                          const float *weights, const float *fm_hessian, const int32_t *pix, int32_t npix,
                          double fm_constant, const mbx_qggmrf *qparams, const float *prox_input, double sigma_prox,
                          int32_t positivity, int32_t delta_dtype, double *stats, void *workspace,
                          int64_t workspace_bytes, void *stream) {
    return vcd_update(g, angles, flat_recon, e, e_dtype, weights, fm_hessian, pix, npix, fm_constant, qparams,
                      prox_input, sigma_prox, positivity, delta_dtype, nullptr, 0, stats, workspace, workspace_bytes,
                      stream);
}

int mbx_vcd_subset_update_planned(const mbx_geom *g, const float *angles, float *flat_recon, void *e,
                                  int32_t e_dtype, const float *weights, const float *fm_hessian,
                                  const int32_t *pix, int32_t npix, double fm_constant, const mbx_qggmrf *qparams,
                                  const float *prox_input, double sigma_prox, int32_t positivity,
                                  int32_t delta_dtype, void *plan, int64_t plan_bytes, double *stats,
                                  void *workspace, int64_t workspace_bytes, void *stream) {
    if (!plan) return fail(MBX_EINVAL, "NULL plan");
    return vcd_update(g, angles, flat_recon, e, e_dtype, weights, fm_hessian, pix, npix, fm_constant, qparams,
                      prox_input, sigma_prox, positivity, delta_dtype, plan, plan_bytes, stats, workspace,
                      workspace_bytes, stream);
}

}  // extern "C"
