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
#include "svdi_internal.hpp"

namespace svdi {
namespace {

constexpr int kThreads = 256;

struct Footprint {
    int a0, na, b0, nb;     // first grid row / column (array coordinates) and extents (<= 9)
    float da, db;           // a0 - p1 and b0 - p2 (centred coordinates), cells
    int local_slab;
    float s;                // Chebyshev argument in [-1, 1]
    float2 vw;              // visibility / |det|
};

template <bool kSnap>
__device__ __forceinline__ bool setup_sample(const SampleInput& in, const GridParams& gp,
                                             size_t idx, size_t row, size_t chan, Footprint& f) {
    const long gmin = -static_cast<long>(gp.G) / 2;
    const long gmax = (static_cast<long>(gp.G) - 1) / 2;
    const long off = static_cast<long>(gp.G) / 2;
    float2 v;
    long lo1, hi1, lo2, hi2;
    float da, db, s;
    int slab = 0;
    if constexpr (kSnap) {
        const double sc = in.freq[chan] / kSpeedOfLight;
        const double p1 = in.rot[3 * row] * sc * gp.scale_f64;
        const double p2 = in.rot[3 * row + 1] * sc * gp.scale_f64;
        const double r3 = in.rot[3 * row + 2] * sc;
        double centre;
        if (gp.archive) {
            centre = gp.r3_centre;
        } else {
            slab = static_cast<int>(floor((r3 - static_cast<double>(gp.r3_min)) / gp.slab_width));
            slab = min(max(slab, 0), gp.total_slabs - 1);
            if (slab < gp.slab_start || slab >= gp.slab_start + gp.nslabs) return false;
            centre = gp.r3_min + (static_cast<float>(slab) + 0.5f) * gp.slab_width;
        }
        s = static_cast<float>(fmin(1.0, fmax(-1.0, (r3 - centre) * static_cast<double>(gp.inv_half))));
        lo1 = max(static_cast<long>(ceil(p1 - 4.0)), gmin);
        hi1 = min(static_cast<long>(floor(p1 + 4.0)), gmax);
        lo2 = max(static_cast<long>(ceil(p2 - 4.0)), gmin);
        hi2 = min(static_cast<long>(floor(p2 + 4.0)), gmax);
        da = static_cast<float>(static_cast<double>(lo1) - p1);
        db = static_cast<float>(static_cast<double>(lo2) - p2);
        const float w0 = in.f0[idx] ? 0.0f : (in.w0 ? in.w0[idx] : 1.0f);
        const float w3 = in.f3[idx] ? 0.0f : (in.w3 ? in.w3[idx] : 1.0f);
        const float ws = w0 + w3;
        if (ws > 0.0f) {
            v.x = (in.v0r[idx] * w0 + in.v3r[idx] * w3) / ws;
            v.y = (in.v0i[idx] * w0 + in.v3i[idx] * w3) / ws;
        } else {
            return false;                       // collapsed visibility is exactly zero
        }
    } else {
        const float p1 = in.bin[3 * idx] * gp.scale_f32;
        const float p2 = in.bin[3 * idx + 1] * gp.scale_f32;
        const float r3 = in.bin[3 * idx + 2];
        float centre;
        if (gp.archive) {
            centre = gp.r3_centre;
        } else {
            slab = static_cast<int>(floorf((r3 - gp.r3_min) / gp.slab_width));
            slab = min(max(slab, 0), gp.total_slabs - 1);
            if (slab < gp.slab_start || slab >= gp.slab_start + gp.nslabs) return false;
            centre = gp.r3_min + (static_cast<float>(slab) + 0.5f) * gp.slab_width;
        }
        s = fminf(1.0f, fmaxf(-1.0f, (r3 - centre) * gp.inv_half));
        lo1 = max(static_cast<long>(ceilf(p1 - 4)), gmin);
        hi1 = min(static_cast<long>(floorf(p1 + 4)), gmax);
        lo2 = max(static_cast<long>(ceilf(p2 - 4)), gmin);
        hi2 = min(static_cast<long>(floorf(p2 + 4)), gmax);
        da = static_cast<float>(lo1) - p1;
        db = static_cast<float>(lo2) - p2;
        v.x = in.vis_re[idx];
        v.y = in.vis_im[idx];
        if (v.x == 0.0f && v.y == 0.0f) return false;
    }
    // Non-finite coordinates convert to arbitrary integers: reject them, and any footprint
    // wider than the 9 cells a finite position can cover.
    if (!(fabsf(da) <= 5.0f) || !(fabsf(db) <= 5.0f)) return false;
    if (lo1 > hi1 || lo2 > hi2 || hi1 - lo1 > 8 || hi2 - lo2 > 8) return false;
    f.a0 = static_cast<int>(lo1 + off);
    f.na = static_cast<int>(hi1 - lo1 + 1);
    f.b0 = static_cast<int>(lo2 + off);
    f.nb = static_cast<int>(hi2 - lo2 + 1);
    f.da = da;
    f.db = db;
    f.local_slab = slab - (gp.archive ? 0 : gp.slab_start);
    f.s = s;
    f.vw.x = v.x / gp.det;
    f.vw.y = v.y / gp.det;
    return true;
}

// One thread per sample; every (cell, term) contribution is a device-scope FP32 atomic add.
template <bool kSnap>
__global__ void __launch_bounds__(kThreads)
k_grid_atomic(SampleInput in, GridParams gp, float2* __restrict__ planes) {
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

size_t launch_grid_tiled(svdi_context*, const SampleInput&, const GridParams&, float2*, int) {
    throw UnsupportedError("tiled gridder not built");
}

void launch_grid_atomic(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                        float2* planes) {
    const size_t plane_bytes = static_cast<size_t>(gp.G) * gp.G * sizeof(float2);
    launch_zero(ctx, planes, plane_bytes * static_cast<size_t>(gp.nslabs) * gp.K);
    if (in.from_snapshot) {
        dim3 grid(static_cast<unsigned>((in.num_rows + kThreads - 1) / kThreads),
                  static_cast<unsigned>(in.num_channels));
        k_grid_atomic<true><<<grid, kThreads, 0, ctx->stream>>>(in, gp, planes);
    } else {
        const unsigned blocks = static_cast<unsigned>((in.num_samples + kThreads - 1) / kThreads);
        k_grid_atomic<false><<<blocks, kThreads, 0, ctx->stream>>>(in, gp, planes);
    }
    SVDI_LAUNCH_CHECK();
}

}  // namespace svdi
