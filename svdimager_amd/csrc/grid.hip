// ES gridding of Chebyshev w-moments (fused_gridding, SVDIntegratedImager.cu:1218-1293;
// gridding, archive/FIPkernels.cu:391-452).
//
// Sample model shared by every gridder: position p = r * cell * G in grid cells (r1 -> grid
// row, r2 -> grid column), footprint ceil(p - 4) .. floor(p + 4) clipped to the grid,
// weight ES((g - p) / 4) per axis, checkerboard (-1)^(g1 + g2) (an fftshift of the output),
// Chebyshev moment T_k(s) with s the sample's r3 scaled into its slab, visibility / |det|.
// Plane index = local_slab * K + k, cell index = (g1 + G/2) * G + (g2 + G/2).
//
// Inputs come either preprocessed (Bin / vis, FP32, as FIpipe receives them) or as a raw
// snapshot: then r = rot_row * nu / c is formed in FP64 here and XX / YY are collapsed with
// their flags and weights on the fly (collapse_visibility_kernel, SVDIntegratedImager.cu:588-617).
#include "sample.hpp"

namespace svdi {
namespace {

constexpr int kThreads = 256;

// One thread per sample; every (cell, term) contribution is a device-scope FP32 atomic add.
template <bool kSnap>
__global__ void __launch_bounds__(kThreads)
k_grid_atomic(SampleInput in, GridParams gp, float2* __restrict__ planes) {
    take_device_plan(gp);
    size_t idx, row = 0, chan = 0;
    if constexpr (kSnap) {
        row = blockIdx.x * (size_t)kThreads + threadIdx.x;
        chan = blockIdx.y;
        if (row >= in.num_rows) return;
        idx = row + chan * in.num_rows;
    } else {
        idx = blockIdx.x * (size_t)kThreads + threadIdx.x;
        if (idx >= in.num_samples) return;
    }
    Footprint f;
    if (!setup_sample<kSnap>(in, gp, idx, row, chan, f)) return;
    float cheb[kMaxTerms];
    chebyshev_sequence(f.s, gp.K, cheb);
    float k1[9], k2[9];
    for (int i = 0; i < f.na; ++i) k1[i] = es_kernel((f.da + i) * 0.25f);
    for (int j = 0; j < f.nb; ++j) k2[j] = es_kernel((f.db + j) * 0.25f);
    const size_t plane_size = static_cast<size_t>(gp.G) * gp.G;
    float2* base = planes + static_cast<size_t>(f.local_slab) * gp.K * plane_size;
    for (int i = 0; i < f.na; ++i) {
        const int a = f.a0 + i;
        for (int j = 0; j < f.nb; ++j) {
            const int b = f.b0 + j;
            float kv = k1[i] * k2[j];
            if ((a + b) & 1) kv = -kv;       // parity of g1 + g2 equals parity of a + b
            float2* cell = base + static_cast<size_t>(a) * gp.G + b;
            for (int k = 0; k < gp.K; ++k) {
                const float m = cheb[k] * kv;
                float* c = reinterpret_cast<float*>(cell + k * plane_size);
                unsafeAtomicAdd(c, f.vw.x * m);
                unsafeAtomicAdd(c + 1, f.vw.y * m);
            }
        }
    }
}

__global__ void k_fill(float* __restrict__ p, size_t n, float value) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_zero(float4* __restrict__ p, size_t n4) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

}  // namespace

void launch_zero(svdi_context* ctx, void* p, size_t bytes) {
    if (bytes % 16 != 0) throw ArgError("launch_zero: size must be a multiple of 16 bytes");
    const size_t n4 = bytes / 16;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n4 + 255) / 256, 8192));
    k_zero<<<blocks, 256, 0, ctx->stream>>>(static_cast<float4*>(p), n4);
    SVDI_LAUNCH_CHECK();
}

void launch_fill(svdi_context* ctx, float* p, size_t count, float value) {
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((count + 255) / 256, 8192));
    if (count == 0) return;
    k_fill<<<blocks, 256, 0, ctx->stream>>>(p, count, value);
    SVDI_LAUNCH_CHECK();
}

void launch_grid_atomic(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                        float2* planes) {
    const size_t plane_bytes = static_cast<size_t>(gp.G) * gp.G * sizeof(float2);
    launch_zero(ctx, planes, plane_bytes * static_cast<size_t>(gp.nslabs) * gp.K);
    const size_t tok = stage_begin(ctx, 5);
    if (in.from_snapshot) {
        dim3 grid(static_cast<unsigned>((in.num_rows + kThreads - 1) / kThreads),
                  static_cast<unsigned>(in.num_channels));
        k_grid_atomic<true><<<grid, kThreads, 0, ctx->stream>>>(in, gp, planes);
    } else {
        const unsigned blocks = static_cast<unsigned>((in.num_samples + kThreads - 1) / kThreads);
        k_grid_atomic<false><<<blocks, kThreads, 0, ctx->stream>>>(in, gp, planes);
    }
    SVDI_LAUNCH_CHECK();
    stage_end(ctx, tok);
}

}  // namespace svdi
