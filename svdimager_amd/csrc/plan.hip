// Chebyshev / slab planning on the device (SVDIntegratedImager.cu:926-1202, 1519-1603;
// archive/FIPkernels.cu:210-366).
//
// r3 range and t3 statistics are block-reduced, finalised by one block that also applies the
// slab rules, and the sampled Chebyshev truncation error is evaluated in FP64 for every
// candidate K in the same launch sequence, so the host reads back a single plan record.
#include "svdi_internal.hpp"

namespace svdi {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ void atomic_max_nonneg(uint32_t* addr, float v) {
    atomicMax(addr, __float_as_uint(v));   // ordering of non-negative floats matches their bits
}

// r3 range of a preprocessed Bin (r3_range_blocks_kernel, SVDIntegratedImager.cu:993-1027).
__global__ void __launch_bounds__(kThreads)
k_bin_r3_range(const float* __restrict__ bin, size_t n, float* __restrict__ partials) {
    __shared__ float lo[kThreads], hi[kThreads];
    float a = INFINITY, b = -INFINITY;
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float r3 = bin[3 * i + 2];
        a = fminf(a, r3);
        b = fmaxf(b, r3);
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            lo[threadIdx.x] = fminf(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmaxf(hi[threadIdx.x], hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = lo[0];
        partials[2 * blockIdx.x + 1] = hi[0];
    }
}

// max |t3| and the count of pixels with a real t3 (t3_stats_blocks_kernel,
// SVDIntegratedImager.cu:1029-1077), FP32 as the reference plans with it.
__global__ void __launch_bounds__(kThreads)
k_t3_stats(const float* __restrict__ vin, size_t N, float cell, float* __restrict__ pmax,
           uint32_t* __restrict__ pvalid) {
    __shared__ float mx[kThreads];
    __shared__ uint32_t cnt[kThreads];
    const size_t total = N * N;
    const float half = static_cast<float>(N) / 2.0f;
    const float v13 = vin[2], v23 = vin[5], v33 = vin[8];
    float m = 0.0f;
    uint32_t c = 0;
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < total; p += (size_t)gridDim.x * kThreads) {
        const size_t y = p / N, x = p - y * N;
        const float t1 = cell * (static_cast<float>(y) - half);
        const float t2 = cell * (static_cast<float>(x) - half);
        const float ra = 1.0f - (t1 + v13) * (t1 + v13) - (t2 + v23) * (t2 + v23);
        if (ra >= 0.0f) {
            const float bs = (v33 >= 0.0f) ? 1.0f : -1.0f;
            m = fmaxf(m, fabsf(-v33 + bs * sqrtf(ra)));
            ++c;
        }
    }
    mx[threadIdx.x] = m;
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            mx[threadIdx.x] = fmaxf(mx[threadIdx.x], mx[threadIdx.x + s]);
            cnt[threadIdx.x] += cnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = mx[0];
        pvalid[blockIdx.x] = cnt[0];
    }
}

// Finalise: r3 range, t3 stats, slab configuration (and the rotation rows, read back with them).
// Integrated: finish_host_chebyshev_selection + make_chebyshev_slab_config (SVDIntegratedImager.cu:941-991).
// Archive: one slab centred on the r3 range (archive/FIPkernels.cu:336-358).
__global__ void __launch_bounds__(kThreads)
k_plan_finalize(const float* __restrict__ r3p, int r3_blocks, const float* __restrict__ t3max,
                const uint32_t* __restrict__ t3valid, int t3_blocks, int archive,
                float target_alpha, const float* __restrict__ vin, DevPlan* __restrict__ plan) {
    __shared__ float lo[kThreads], hi[kThreads], mx[kThreads];
    __shared__ uint32_t cnt[kThreads];
    float a = INFINITY, b = -INFINITY, m = 0.0f;
    uint32_t c = 0;
    for (int i = threadIdx.x; i < r3_blocks; i += kThreads) {
        a = fminf(a, r3p[2 * i]);
        b = fmaxf(b, r3p[2 * i + 1]);
    }
    for (int i = threadIdx.x; i < t3_blocks; i += kThreads) {
        m = fmaxf(m, t3max[i]);
        c += t3valid[i];
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    mx[threadIdx.x] = m;
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            lo[threadIdx.x] = fminf(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmaxf(hi[threadIdx.x], hi[threadIdx.x + s]);
            mx[threadIdx.x] = fmaxf(mx[threadIdx.x], mx[threadIdx.x + s]);
            cnt[threadIdx.x] += cnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const float r3_min = lo[0], r3_max = hi[0], max_t3 = mx[0];
    DevPlan p = {};
    p.r3_min = r3_min;
    p.r3_max = r3_max;
    p.max_abs_t3 = max_t3;
    p.valid_pixels = cnt[0];
    p.r3_centre = 0.5f * (r3_max + r3_min);
    p.r3_half_range = fmaxf(0.5f * (r3_max - r3_min), 1.0e-6f);
    // 2.0f * M_PI * ... evaluates in double in the reference (M_PI is a double literal).
    p.alpha_max = static_cast<float>(2.0f * kPi * max_t3 * p.r3_half_range);
    p.target_alpha = target_alpha;
    if (archive) {
        p.num_slabs = 1;
        p.r3_range = r3_max - r3_min;
        p.slab_width = p.r3_range;
        p.slab_half_range = p.r3_half_range;
        p.inv_slab_half_range = 1.0f / p.r3_half_range;
        p.slab_alpha_max = p.alpha_max;
    } else {
        p.r3_range = fmaxf(r3_max - r3_min, 2.0e-6f);
        int ns = static_cast<int>(ceilf(fmaxf(p.alpha_max, 0.0f) / target_alpha));
        if (ns < 1) ns = 1;
        p.num_slabs = ns;
        p.slab_width = p.r3_range / static_cast<float>(ns);
        p.slab_half_range = fmaxf(0.5f * p.slab_width, 1.0e-6f);
        p.inv_slab_half_range = 1.0f / p.slab_half_range;
        p.slab_alpha_max = static_cast<float>(2.0f * kPi * max_t3 * p.slab_half_range);
    }
    for (int i = 0; i < 9; ++i) p.vin[i] = vin[i];
    p.inv_slab_width = 1.0 / static_cast<double>(p.slab_width);
    // |v0 v4 - v1 v3| with each operation rounded on its own, as the host forms it.
    p.det = fabsf(__fsub_rn(__fmul_rn(vin[0], vin[4]), __fmul_rn(vin[1], vin[3])));
    p.inv_det = 1.0f / p.det;
    *plan = p;
}

// Sampled max |e^{i a s} - sum_{n<K} (n ? 2 : 1) i^n J_n(a) T_n(s)| over a in [0, alpha_max]
// and s in [-1, 1], evaluated in FP64.  One block per alpha sample, two or four s samples per
// thread: every thread forms its e^{i a s} while the Bessel sequence is formed (for a < 2 by the
// power series, one order per lane of the first wavefront; otherwise by thread 0's recurrence).
//   integrated: 257 x 1024 samples, K in {8, 16, 24, 32} (cheb_bucket_error_kernel,
//               SVDIntegratedImager.cu:1125-1202; alpha and s formed in FP32 as there);
//   archive:    513 x 2048 samples, every K in 1..32 (host_chebyshev_error_sample_range,
//               archive/FIPkernels.cu:272-318; alpha in FP32, s in FP64 as there).
constexpr int kTermThreads = 512;
constexpr double kBesselLaneSeriesMax = 2.0;

template <bool kArchive>
__global__ void __launch_bounds__(kTermThreads)
k_term_errors(DevPlan* __restrict__ plan) {
    constexpr int na = kArchive ? 513 : 257;
    constexpr int ns = kArchive ? 2048 : 1024;
    constexpr int kPer = ns / kTermThreads;     // s samples per thread
    // Errors kept: every K (archive) or the buckets K = 8, 16, 24, 32 (terms 7, 15, 23, 31).
    constexpr int kKept = kArchive ? kMaxTerms : kMaxTerms / 8;
    __shared__ double jn[kMaxTerms];
    __shared__ double red[kKept][kTermThreads / 64];
    const float amax = plan->slab_alpha_max;
    const float alpha_f = amax * static_cast<float>(blockIdx.x) / static_cast<float>(na - 1);
    const double alpha = static_cast<double>(alpha_f);
    double sv[kPer], er[kPer], ei[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int si = threadIdx.x + u * kTermThreads;
        if constexpr (kArchive) sv[u] = -1.0 + 2.0 * static_cast<double>(si) / static_cast<double>(ns - 1);
        else sv[u] = static_cast<double>(-1.0f + 2.0f * static_cast<float>(si) / static_cast<float>(ns - 1));
        sincos(alpha * sv[u], &ei[u], &er[u]);
    }
    if (fabs(alpha) < kBesselLaneSeriesMax) {
        if (threadIdx.x < kMaxTerms) jn[threadIdx.x] = bessel_series(static_cast<int>(threadIdx.x), alpha);
    } else if (threadIdx.x == 0) {
        double j[kMaxTerms];
        bessel_sequence_fixed<kMaxTerms>(alpha, j);
#pragma unroll
        for (int n = 0; n < kMaxTerms; ++n) jn[n] = j[n];
    }
    __syncthreads();
    double err2[kKept];                      // squared error; the root is taken once at the end
#pragma unroll
    for (int i = 0; i < kKept; ++i) err2[i] = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const double s = sv[u];
        double ar = 0.0, ai = 0.0, tm2 = 1.0, tm1 = s;
#pragma unroll
        for (int n = 0; n < kMaxTerms; ++n) {
            double tn;
            if (n == 0) tn = 1.0;
            else if (n == 1) tn = s;
            else {
                tn = 2.0 * s * tm1 - tm2;
                tm2 = tm1;
                tm1 = tn;
            }
            if (n == 0) {
                ar += jn[0] * tn;
            } else {
                const double amp = 2.0 * jn[n] * tn;
                switch (n & 3) {
                    case 0: ar += amp; break;
                    case 1: ai += amp; break;
                    case 2: ar -= amp; break;
                    default: ai -= amp; break;
                }
            }
            if (kArchive || (n & 7) == 7) {
                const int i = kArchive ? n : n >> 3;
                const double dr = er[u] - ar, di = ei[u] - ai;
                err2[i] = fmax(err2[i], dr * dr + di * di);
            }
        }
    }
    // Maxima of the squared errors: per wavefront, then across the block's wavefronts, then one
    // root per kept term count.
#pragma unroll
    for (int i = 0; i < kKept; ++i) {
        double v = err2[i];
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
        if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < kKept) {
        double v2 = 0.0;
        for (int w = 0; w < kTermThreads / 64; ++w) v2 = fmax(v2, red[threadIdx.x][w]);
        const float v = static_cast<float>(sqrt(v2));
        if constexpr (kArchive) atomic_max_nonneg(&plan->term_err_bits[threadIdx.x], v);
        else atomic_max_nonneg(&plan->bucket_err_bits[threadIdx.x], v);
    }
}

}  // namespace

void launch_plan(svdi_context* ctx, const SampleInput& in, Semantics sem, size_t image_size,
                 float cell, float target_alpha, DevPlan* plan) {
    const int archive = sem == Semantics::Archive ? 1 : 0;
    // r3 partials: from the row rotation (snapshot input) or a reduction over Bin.
    float* r3p;
    int r3_blocks;
    char* misc = static_cast<char*>(ctx->ws.get(Workspace::kMisc, 1 << 20));
    if (in.r3_partials != nullptr) {
        r3p = const_cast<float*>(in.r3_partials);
        r3_blocks = in.r3_blocks;
    } else {
        r3_blocks = 1024;
        r3p = reinterpret_cast<float*>(misc);
        k_bin_r3_range<<<r3_blocks, kThreads, 0, ctx->stream>>>(in.bin, in.num_samples, r3p);
        SVDI_LAUNCH_CHECK();
    }
    const int t3_blocks = 512;
    float* t3max = reinterpret_cast<float*>(misc + 64 * 1024);
    uint32_t* t3valid = reinterpret_cast<uint32_t*>(misc + 128 * 1024);
    const float* vin = in.from_snapshot ? in.geo->vin : in.vin_dev;
    k_t3_stats<<<t3_blocks, kThreads, 0, ctx->stream>>>(vin, image_size, cell, t3max, t3valid);
    SVDI_LAUNCH_CHECK();
    k_plan_finalize<<<1, kThreads, 0, ctx->stream>>>(r3p, r3_blocks, t3max, t3valid, t3_blocks,
                                                    archive, target_alpha, vin, plan);
    SVDI_LAUNCH_CHECK();
    if (archive) k_term_errors<true><<<513, kTermThreads, 0, ctx->stream>>>(plan);
    else k_term_errors<false><<<257, kTermThreads, 0, ctx->stream>>>(plan);
    SVDI_LAUNCH_CHECK();
}

}  // namespace svdi
