// Exclusive scan of a u32 array on the device (counts -> offsets), shared by the binning passes
// of the gridders: per-block scans of 4096 elements, a one-block scan of the block totals, then
// the block offsets added; out[n] receives the total and `cursor` (optional) a copy of out.
#pragma once

#include "svdi_internal.hpp"

namespace svdi {
namespace {

constexpr int kScanPer = 4;              // elements per thread
constexpr int kScanBlock = 1024 * kScanPer;

// Block-wide inclusive scan of one value per thread (blockDim.x = 1024).
__device__ __forceinline__ uint32_t block_scan_1024(uint32_t v, uint32_t* s_wave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t w = threadIdx.x < 16 ? s_wave[threadIdx.x] : 0u;
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t t = __shfl_up(w, d, 64);
            if (lane >= d) w += t;
        }
        if (threadIdx.x < 16) s_wave[threadIdx.x] = w;
    }
    __syncthreads();
    const uint32_t r = v + (wave > 0 ? s_wave[wave - 1] : 0u);
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(1024)
k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t* __restrict__ sums, uint32_t n) {
    __shared__ uint32_t s_wave[16];
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanBlock + threadIdx.x * kScanPer;
    uint32_t v[kScanPer], tot = 0;
    for (int j = 0; j < kScanPer; ++j) {
        v[j] = base + j < n ? in[base + j] : 0u;
        tot += v[j];
    }
    const uint32_t incl = block_scan_1024(tot, s_wave);
    uint32_t run = incl - tot;
    for (int j = 0; j < kScanPer; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 1023) sums[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(1024)
k_scan_sums(uint32_t* __restrict__ sums, int nblocks, uint32_t* __restrict__ grand_total) {
    __shared__ uint32_t s_wave[16];
    const int per = (nblocks + 1023) / 1024;
    const int lo = min(static_cast<int>(threadIdx.x) * per, nblocks), hi = min(lo + per, nblocks);
    uint32_t tot = 0;
    for (int i = lo; i < hi; ++i) tot += sums[i];
    const uint32_t incl = block_scan_1024(tot, s_wave);
    uint32_t run = incl - tot;
    for (int i = lo; i < hi; ++i) {
        const uint32_t t = sums[i];
        sums[i] = run;
        run += t;
    }
    if (threadIdx.x == 1023) *grand_total = incl;
}

__global__ void __launch_bounds__(256)
k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ sums, uint32_t n, uint32_t* __restrict__ cursor) {
    const size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = out[i] + sums[i / kScanBlock];
    out[i] = v;
    if (cursor != nullptr) cursor[i] = v;
}

inline size_t scan_blocks(size_t n) { return (n + kScanBlock - 1) / kScanBlock; }

// sums: scan_blocks(n) elements of scratch; out: n + 1 elements.
inline void launch_exclusive_scan(svdi_context* ctx, const uint32_t* in, uint32_t* out, uint32_t* cursor,
                                  uint32_t* sums, size_t n) {
    if (n >= (1ull << 32)) throw UnsupportedError("scan too large");
    const size_t nb = scan_blocks(n);
    if (nb > (1ull << 30)) throw UnsupportedError("scan too large");
    const uint32_t n32 = static_cast<uint32_t>(n);
    k_scan_blocks<<<static_cast<unsigned>(nb), 1024, 0, ctx->stream>>>(in, out, sums, n32);
    SVDI_LAUNCH_CHECK();
    k_scan_sums<<<1, 1024, 0, ctx->stream>>>(sums, static_cast<int>(nb), out + n);
    SVDI_LAUNCH_CHECK();
    k_scan_add<<<static_cast<unsigned>((n + 255) / 256), 256, 0, ctx->stream>>>(out, sums, n32, cursor);
    SVDI_LAUNCH_CHECK();
}

}  // namespace
}  // namespace svdi
