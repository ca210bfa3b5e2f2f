// MFMA gridder: the ES gridding of Chebyshev moments (fused_gridding, SVDIntegratedImager.cu:
// 1218-1293; gridding, archive/FIPkernels.cu:391-452) recast as dense contractions over
// output-stationary 16 x 16 windows of the grid.
//
// For a window W and the samples r whose footprint meets it, every (term k, re/im, cell) of W is
//   D[m][(a, b)] = sum_r c[m][r] * ku_r[a] * kv_r[b],   m = 2k + ri,
//   c[2k + ri][r] = T_k(s_r) * (Re | Im) v_r / |det|,   ku_r[a] = ES((a - p1_r) / 4) (0 off support),
// with (a, b) the window's cells and p_r the sample's position relative to the window.  For each
// window row a, D_a = C * B_a runs on v_mfma_f32_16x16x32_f16, with C = c (16 x 32 samples) and
// B_a[r][b] = ku_r[a] kv_r[b] (32 samples x 16 columns), the 16 row blocks (64 FP32 accumulators
// per lane) staying in registers across the window's samples.  A window is written once: by the
// wavefront or workgroup that contracted it, or, when it was split over several workgroup
// pieces, as partial windows that k_win_reduce adds in piece order; no atomics touch the planes.
// Precision: c and the ES factors are split into f16 hi + lo and each product taken as
// hi*hi + lo*hi + hi*lo with FP32 accumulation (~22 significant bits per product); c is scaled by
// a power of two fixed from the largest |v| of the call and the ES factors by 2^7 each, so
// nothing overflows f16; the scales are undone exactly at write-out.
//
// Binning by row class: with class height c (4, or 8 when the finer histogram does not fit in
// LDS), a sample whose (clipped) first footprint row a0 lies in [c h, c h + c) is in row class h,
// and its footprint rows lie in [c h, c h + c + 8).  Window row wa (rows [16 wa, 16 wa + 16))
// therefore reads the 24 / c classes h = (16 / c) wa + d, d = -8 / c .. 16 / c - 1, class d
// touching only the window rows [c d, c d + c + 7].  Records are binned by (slab, window
// column, row class) with the class fastest, so a window's records are one contiguous range
// ordered by class, each record is written once per window column it meets, and a chunk of 32
// records drawn from neighbouring classes skips the 4-row quarters of the window that none of
// them touches (with c = 4, 12 of 16 rows per class).  Counting uses block-local
// LDS histograms kept per block; a prefix over the blocks and a scan of the bin totals give each
// block its first slot in every bin, and one more pass over its samples writes the records
// (position relative to the class's home window row / the window column, s, v).  When the bins
// do not fit in one LDS histogram they are split into bands of consecutive bins, and each band
// is a separate set of workgroups that reads the block's samples and handles only its own bins.
#include "sample.hpp"
#include "scan.hpp"
#include "es_poly.hpp"
#include "records.hpp"

#include <cstring>

namespace svdi {
namespace {

constexpr int kWin = 16;                 // window side in cells
constexpr int kMaxClasses = 6;           // classes per window: 24 / class height (4 or 8)
constexpr int kBinThreads = 1024;        // binning blocks: 256 rows x 4 channel groups at a time
constexpr int kBinRowSets = 2;           // row sets of 256 per block (fewer, larger blocks: the
                                         // per-block histograms are written and read back whole)
constexpr int kBinRowsPerSet = 256;
constexpr int kBinRows = kBinRowSets * kBinRowsPerSet;
constexpr int kBinChanGroups = kBinThreads / kBinRowsPerSet;
constexpr int kBinPerThread = 16;        // samples per thread, preprocessed input
constexpr int kLdsBins = 36864;          // largest block-local histogram (144 KB of LDS)
constexpr size_t kMaxBlockHistBytes = size_t(1) << 31;   // per-block histograms of all bands
constexpr int kChunk = 32;               // samples per MFMA reduction step
constexpr int kStride = 36;              // floats per c staging row (32 samples + pad)
constexpr int kStrideH = 40;             // halves per ES staging row (32 samples + pad, 16-byte rows)
constexpr int kPad = 8;                  // ES staging rows beyond each window edge (-8 .. 23)
constexpr int kStageRows = kWin + 2 * kPad;
constexpr int kPlaneHalves = kStageRows * kStrideH;               // one ES plane (hi or lo)
constexpr int kEsScaleLog2Sq = 14;      // ES factors scaled by 2^7 per axis (in the polynomial table)

// Operand precision of the window contraction D_a += C B_a (C = c, B_a = ku_a kv).
enum WinPrec : int {
    kPrecSplit3 = 0,   // c and both ES factors in f16 hi + lo; B_a's product error kept:
                       // hi*hi + lo*hi + hi*lo, ~22 significant bits per product, 3 MFMAs per row
    kPrecSplitC = 1,   // c in f16 hi + lo, B_a = f16(f16(ku) f16(kv)): 2 MFMAs per row
    kPrecSingle = 2,   // c and B_a single f16: 1 MFMA per row
    kPrecBf16 = 3,     // c and B_a in bf16 hi + lo, B_a's product formed in FP32: hi*hi + lo*hi +
                       // hi*lo on the bf16 MFMA, ~16 significant bits per product
};
constexpr int kStrideF = 36;                               // floats per FP32 ES staging row
constexpr int kPlaneFloats = kStageRows * kStrideF;        // one FP32 ES plane
template <int P>
struct WinTraits {
    static constexpr bool kBf16 = P == kPrecBf16;          // FP32 ES planes, bf16 operands
    static constexpr bool kEsLo = P == kPrecSplit3;        // ES factors staged as f16 hi + lo planes
    static constexpr bool kCLo = P != kPrecSingle;         // c split into hi + lo
    static constexpr bool kDither = P == kPrecSplitC || P == kPrecSingle;   // single-f16 ES product
    static constexpr int kPlanes = kEsLo ? 4 : 2;
    // Per wavefront: the ES planes (ku hi [lo], kv hi [lo], f16; or ku, kv in FP32), then c[16]
    // (f32).
    static constexpr int kStageFloats =
        (kBf16 ? 2 * kPlaneFloats : (kPlanes * kPlaneHalves) / 2) + kWin * kStride;
    // Wavefronts per workgroup (one workgroup per CU): as many as the staging allows.  At 12 the
    // kernel keeps its ~150 registers without spills; 16 (four per SIMD) would need 128 and a
    // staging area without row padding, which measured slower (542 against 500 us for split3).
    static constexpr int kWaves = (kEsLo || kBf16) ? 12 : 16;
    // Accumulator slots of the workgroup reduction (within the staging area).
    static constexpr int kSlots = (kEsLo || kBf16) ? 6 : 4;
    static_assert(kWaves % kSlots == 0, "reduction rounds");
    static_assert(kSlots * kWin * 64 * 4 <= kStageFloats * kWaves, "reduction slots exceed the staging area");
};
// ES(1) 2^7, the ninth footprint cell of an integer position, as f16 hi + lo.
constexpr _Float16 kEsPolyEdgeHi = static_cast<_Float16>(kEsPolyEdge);
constexpr _Float16 kEsPolyEdgeLo = static_cast<_Float16>(kEsPolyEdge - static_cast<float>(kEsPolyEdgeHi));

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

struct WinGeom {
    int nw;          // windows per grid axis
    int cls_log2;    // log2 of the row-class height (2 or 3)
    int ncls;        // classes per window: 24 >> cls_log2
    int nh;          // row classes: ceil(G / class height)
    int nwin;        // nw * nw
    int nwin_total;  // nslabs * nwin
    int nbins;       // nslabs * nw * nh, ordered (slab, window column, row class)
    int use_lds;     // block histograms in LDS (else one global atomic per record)
    int band_bins;   // bins per LDS band (<= kLdsBins); band y covers [y band_bins, ...)
    uint32_t coop_min;   // records up to which a window is one wavefront item
    uint32_t coop_cap;   // records per workgroup piece
};

// A work item: records [start, start + count) of window `win`; bnd[j] (j = 1 .. ncls - 1) is the
// first record of the window's class j (d = j - 8 / c), records before bnd[1] are of class 0,
// and bnd[j] for j >= ncls is the window's end.
// slot >= 0: the item is one piece of a window split over several workgroup items and writes its
// partial window to partial slot `slot`; slot < 0: it writes the window to the planes.
struct WItem {
    int win, start, count, slot;
    int bnd[kMaxClasses];
};

// Records: 16 bytes each (records.hpp), one aligned 16-byte access to write or read.
struct WRecords {
    uint4* r;

    __device__ __forceinline__ void put(uint32_t slot, const RecordWords& w) const {
        r[slot] = make_uint4(w.w[0], w.w[1], w.w[2], w.w[3]);
    }
    __device__ __forceinline__ uint4 get(int i) const { return r[i]; }
};

// Every (sample, bin) pair of this block's samples: one row class h = a0 / c (c the class
// height), and the window columns b0 / 16 .. (b0 + nb - 1) / 16 that the clipped footprint meets
// (at most two).  Snapshot input: 512 consecutive rows in two sets of 256, each row split over 4
// threads that take every fourth channel (a wavefront reads 64 consecutive rows of one channel:
// coalesced), loading four samples before working on the first.  Preprocessed input: 16384
// consecutive samples, 16 per thread at a stride of the block size.  kVis = false (snapshot
// input): neither visibilities nor weights are loaded and f.vw is zero; the bins depend on flags
// and geometry only (samples are kept unless both polarisations are flagged, V = 0 when the
// weights sum to zero), so both passes keep the same samples.
template <bool kSnap, bool kVis, class Fn>
__device__ __forceinline__ void for_sample_bins(const SampleInput& in, const GridParams& gp,
                                                const WinGeom& wg, size_t blk, Fn&& fn) {
    auto bins_of = [&](const Footprint& f) {
        // (a0, b0 >= 0: array coordinates of an in-grid footprint.)
        const int h = f.a0 >> wg.cls_log2;
        const int col0 = f.local_slab * wg.nw;
        const int wb0 = static_cast<int>(static_cast<uint32_t>(f.b0) / kWin);
        const int wb1 = static_cast<int>(static_cast<uint32_t>(f.b0 + f.nb - 1) / kWin);
        for (int wb = wb0; wb <= wb1; ++wb) fn(f, (col0 + wb) * wg.nh + h, h, wb);
    };
    if constexpr (kSnap) {
        if (in.quad_rows) {
            // Four consecutive rows of one channel per thread (one 16-byte load per FP32 array,
            // 4 bytes per flag array; a wavefront reads 256 consecutive rows), 128 row quads x 8
            // channel groups; the next channel's loads are issued before this one is binned.
            constexpr int kQuads = kBinRows / 4;
            constexpr int kQuadGroups = kBinThreads / kQuads;
            const size_t row0 = blk * kBinRows + 4 * (threadIdx.x % kQuads);
            if (row0 >= in.num_rows) return;          // (num_rows % 4 == 0: quads are whole)
            struct QuadLoad {
                double nu;
                float4 v0r, v0i, v3r, v3i, w0, w3;
                uint32_t f0, f3;
            };
            const float4 one4 = make_float4(1.0f, 1.0f, 1.0f, 1.0f), zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            auto load = [&](size_t chan) {
                QuadLoad q;
                const size_t idx = row0 + chan * in.num_rows;
                q.nu = in.freq[chan];
                q.v0r = q.v0i = q.v3r = q.v3i = zero4;
                q.w0 = q.w3 = one4;
                if constexpr (kVis) {
                    q.v0r = *reinterpret_cast<const float4*>(in.v0r + idx);
                    q.v0i = *reinterpret_cast<const float4*>(in.v0i + idx);
                    q.v3r = *reinterpret_cast<const float4*>(in.v3r + idx);
                    q.v3i = *reinterpret_cast<const float4*>(in.v3i + idx);
                    if (in.w0) q.w0 = *reinterpret_cast<const float4*>(in.w0 + idx);
                    if (in.w3) q.w3 = *reinterpret_cast<const float4*>(in.w3 + idx);
                }
                q.f0 = *reinterpret_cast<const uint32_t*>(in.f0 + idx);
                q.f3 = *reinterpret_cast<const uint32_t*>(in.f3 + idx);
                return q;
            };
            auto lane4 = [](const float4& v, int r) { return r == 0 ? v.x : r == 1 ? v.y : r == 2 ? v.z : v.w; };
            size_t chan = threadIdx.x / kQuads;
            if (chan >= in.num_channels) return;
            QuadLoad cur = load(chan);
            double geo[4][3];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 3; ++k) geo[r][k] = in.rot[3 * (row0 + r) + k];
            for (; chan < in.num_channels; chan += kQuadGroups) {
                QuadLoad nxt = cur;
                if (chan + kQuadGroups < in.num_channels) nxt = load(chan + kQuadGroups);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    SnapLoad s;
                    s.nu = cur.nu;
                    s.v0r = lane4(cur.v0r, r);
                    s.v0i = lane4(cur.v0i, r);
                    s.v3r = lane4(cur.v3r, r);
                    s.v3i = lane4(cur.v3i, r);
                    s.w0 = lane4(cur.w0, r);
                    s.w3 = lane4(cur.w3, r);
                    s.f0 = static_cast<uint8_t>(cur.f0 >> (8 * r));
                    s.f3 = static_cast<uint8_t>(cur.f3 >> (8 * r));
                    Footprint f;
                    if (finish_snapshot_sample<true>(gp, geo[r][0], geo[r][1], geo[r][2], s, f)) bins_of(f);
                }
                cur = nxt;
            }
            return;
        }
        constexpr int kBatch = 4;
        for (int set = 0; set < kBinRowSets; ++set) {
        const size_t row = blk * kBinRows + set * kBinRowsPerSet + (threadIdx.x % kBinRowsPerSet);
        if (row >= in.num_rows) return;
        const double ru = in.rot[3 * row], rv = in.rot[3 * row + 1], rw = in.rot[3 * row + 2];
        for (size_t c0 = threadIdx.x / kBinRowsPerSet; c0 < in.num_channels; c0 += kBatch * kBinChanGroups) {
            SnapLoad ld[kBatch] = {};
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const size_t chan = c0 + u * kBinChanGroups;
                if (chan < in.num_channels) ld[u] = load_snapshot_sample<kVis, kVis>(in, row + chan * in.num_rows, chan);
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const size_t chan = c0 + u * kBinChanGroups;
                Footprint f;
                if (chan < in.num_channels && finish_snapshot_sample<true>(gp, ru, rv, rw, ld[u], f)) bins_of(f);
            }
        }
        }
    } else {
        const size_t base = blk * kBinThreads * kBinPerThread + threadIdx.x;
        for (int j = 0; j < kBinPerThread; ++j) {
            const size_t idx = base + static_cast<size_t>(j) * kBinThreads;
            if (idx >= in.num_samples) return;
            Footprint f;
            if (setup_sample<false>(in, gp, idx, 0, 0, f)) bins_of(f);
        }
    }
}

// Largest |Re v|, |Im v| of the block, folded into *vmax_bits (non-negative floats order as
// their bit patterns).
__device__ __forceinline__ void block_max_to(float m, uint32_t* vmax_bits) {
    __shared__ float s_max[kBinThreads / 64];
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = s_max[0];
        for (int w = 1; w < kBinThreads / 64; ++w) b = fmaxf(b, s_max[w]);
        if (b > 0.0f && b <= 3.0e38f) atomicMax(vmax_bits, __float_as_uint(b));
    }
}

// Sum of v over the block's kBinThreads threads, returned to every thread (once per kernel).
__device__ __forceinline__ uint32_t block_sum(uint32_t v) {
    __shared__ uint32_t s_part[kBinThreads / 64];
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < kBinThreads / 64; ++w) t += s_part[w];
    return t;
}

// Bins [lo, lo + n) of LDS band blockIdx.y.
__device__ __forceinline__ void lds_band(const WinGeom& wg, int& lo, int& n) {
    lo = static_cast<int>(blockIdx.y) * wg.band_bins;
    n = min(wg.band_bins, wg.nbins - lo);
}

// Bin counts: per (block, band) in an LDS histogram (dynamic LDS) written to
// blockhist[block * nbins + bin] for the band's bins, and the (block, band)'s record total to
// blocktot[band * blocks + block] (without LDS histograms the counts go straight to counts[bin]).
// kHist16: 16-bit counters, two bins per LDS word (bin b in the half b & 1 of word b / 2, added
// to by a 32-bit atomic of 1 << 16 (b & 1)), for blocks of fewer than 65536 samples: a sample
// adds at most one record to a bin, so no half carries into the other.  Half the LDS lets two
// workgroups share a CU.
template <bool kSnap, bool kHist16, typename CountT>
__global__ void __launch_bounds__(kBinThreads)
k_win_count(SampleInput in, GridParams gp, WinGeom wg, uint32_t* __restrict__ counts,
            CountT* __restrict__ blockhist, uint32_t* __restrict__ blocktot) {
    take_device_plan(gp);
    extern __shared__ uint32_t hist[];
    const size_t blk = blockIdx.x;
    if (!wg.use_lds) {
        for_sample_bins<kSnap, false>(in, gp, wg, blk, [&](const Footprint&, int bin, int, int) {
            atomicAdd(&counts[bin], 1u);
        });
        return;
    }
    int lo, n;
    lds_band(wg, lo, n);
    const int words = kHist16 ? (n + 1) / 2 : n;
    for (int w = threadIdx.x; w < words; w += kBinThreads) hist[w] = 0;
    __syncthreads();
    uint32_t cnt = 0;
    for_sample_bins<kSnap, false>(in, gp, wg, blk, [&](const Footprint&, int bin, int, int) {
        const uint32_t b = static_cast<uint32_t>(bin - lo);
        if (b < static_cast<uint32_t>(n)) {
            if constexpr (kHist16) atomicAdd(&hist[b >> 1], 1u << ((b & 1u) << 4));
            else atomicAdd(&hist[b], 1u);
            ++cnt;
        }
    });
    const uint32_t tot = block_sum(cnt);          // (its barrier also ends the histogram's atomics)
    if (threadIdx.x == 0) blocktot[static_cast<size_t>(blockIdx.y) * gridDim.x + blk] = tot;
    CountT* row = blockhist + blk * wg.nbins + lo;
    for (int b = threadIdx.x; b < n; b += kBinThreads)
        row[b] = static_cast<CountT>(kHist16 ? (hist[b >> 1] >> ((b & 1) << 4)) & 0xffffu : hist[b]);
}

// The per-block histograms turned into per-block record offsets without global atomics per
// (block, bin): the blocks are split into kBinSegs segments; k_bin_seg_sums forms each segment's
// count of every bin (and the bin totals), and after the scan of the totals k_bin_block_offsets
// writes offs[block][bin] = boff[bin] + the records of the bin in earlier blocks.  Each block's
// records of a bin therefore follow those of the blocks before it.  The counts are 16-bit when a
// block has fewer than 65536 samples (half the histogram traffic), and otherwise 32-bit words
// that the offsets then replace in place.
constexpr int kBinSegs = 8;

__device__ __forceinline__ void bin_segment(int nblocks, int seg, int& b0, int& b1) {
    const int per = (nblocks + kBinSegs - 1) / kBinSegs;
    b0 = min(seg * per, nblocks);
    b1 = min(b0 + per, nblocks);
}

template <typename CountT>
__global__ void __launch_bounds__(256)
k_bin_seg_sums(const CountT* __restrict__ blockhist, int nblocks, int nbins, uint32_t* __restrict__ segsum,
               uint32_t* __restrict__ counts) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= kBinSegs * nbins) return;
    const int seg = t / nbins, bin = t - seg * nbins;
    int b0, b1;
    bin_segment(nblocks, seg, b0, b1);
    const CountT* p = blockhist + bin;
    uint32_t sum = 0;
    int b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = p[static_cast<size_t>(b + j) * nbins];
#pragma unroll
        for (int j = 0; j < 8; ++j) sum += v[j];
    }
    for (; b < b1; ++b) sum += p[static_cast<size_t>(b) * nbins];
    segsum[t] = sum;
    if (sum != 0) atomicAdd(&counts[bin], sum);
}

// cnt and offs may be the same words (32-bit counts): each is read before it is written.
template <typename CountT>
__global__ void __launch_bounds__(256)
k_bin_block_offsets(const CountT* cnt, uint32_t* offs, int nblocks, int nbins, const uint32_t* __restrict__ segsum,
                    const uint32_t* __restrict__ boff) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= kBinSegs * nbins) return;
    const int seg = t / nbins, bin = t - seg * nbins;
    int b0, b1;
    bin_segment(nblocks, seg, b0, b1);
    uint32_t run = boff[bin];
    for (int s = 0; s < seg; ++s) run += segsum[s * nbins + bin];
    const CountT* p = cnt + bin;
    uint32_t* o = offs + bin;
    int b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = p[static_cast<size_t>(b + j) * nbins];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            o[static_cast<size_t>(b + j) * nbins] = run;
            run += v[j];
        }
    }
    for (; b < b1; ++b) {
        const uint32_t v = p[static_cast<size_t>(b) * nbins];
        o[static_cast<size_t>(b) * nbins] = run;
        run += v;
    }
}

// Records of this block's samples, in one pass per band: with LDS histograms the block's first
// slot of each of the band's bins comes from blockhist (the offsets of k_bin_block_offsets) and
// LDS cursors hand out the slots; otherwise every record takes a global cursor slot.
// `limit` (the record buffer's capacity) only bounds the writes.  With LDS histograms the
// (block, band)'s emitted records are checked against its count-pass total: a difference (the
// inputs changed between the passes) would leave slots unwritten, so it sets kErrBinMismatch in
// *err and the call fails instead of gridding stale records.
template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_win_scatter(SampleInput in, GridParams gp, WinGeom wg, const uint32_t* __restrict__ blockhist,
              const uint32_t* __restrict__ blocktot, uint32_t* __restrict__ cursor, WRecords rec,
              uint32_t limit, uint32_t* __restrict__ vmax_bits, uint32_t* __restrict__ err) {
    take_device_plan(gp);
    extern __shared__ uint32_t hist[];
    const int off = gp.G / 2;
    float m = 0.0f;
    auto emit = [&](uint32_t slot, const Footprint& f, int h, int wb) {
        m = fmaxf(m, fmaxf(fabsf(f.vw.x), fabsf(f.vw.y)));
        if (slot >= limit) return;
        // Row position relative to the class's home window row (c h + 8) / 16: the row window
        // that reads the class at d < 8 / c (the row above reads it at d >= 8 / c).  Snapshot
        // input: q = p - (first footprint cell); preprocessed input: q = p (centred).
        const int home = ((h << wg.cls_log2) + 8) >> 4;
        const int base1 = kSnap ? f.a0 - home * kWin : off - home * kWin;
        const int base2 = kSnap ? f.b0 - wb * kWin : off - wb * kWin;
        rec.put(slot, rec_encode(base1, f.q1, base2, f.q2, f.s, f.vw.x, f.vw.y));
    };
    const size_t blk = blockIdx.x;
    if (!wg.use_lds) {
        for_sample_bins<kSnap, true>(in, gp, wg, blk, [&](const Footprint& f, int bin, int h, int wb) {
            emit(atomicAdd(&cursor[bin], 1u), f, h, wb);
        });
        block_max_to(m, vmax_bits);
        return;
    }
    int lo, n;
    lds_band(wg, lo, n);
    const uint32_t* row = blockhist + blk * wg.nbins + lo;
    for (int b = threadIdx.x; b < n; b += kBinThreads) hist[b] = row[b];
    __syncthreads();
    uint32_t cnt = 0;
    for_sample_bins<kSnap, true>(in, gp, wg, blk, [&](const Footprint& f, int bin, int h, int wb) {
        const uint32_t b = static_cast<uint32_t>(bin - lo);
        if (b < static_cast<uint32_t>(n)) {
            emit(atomicAdd(&hist[b], 1u), f, h, wb);
            ++cnt;
        }
    });
    const uint32_t tot = block_sum(cnt);
    if (threadIdx.x == 0 && tot != blocktot[static_cast<size_t>(blockIdx.y) * gridDim.x + blk])
        atomicOr(err, kErrBinMismatch);
    block_max_to(m, vmax_bits);
}

// Record range of window w (slab, wa, wb): bnd[j] = first record of the window's class j
// (h = (16 / c) wa - 8 / c + j, clamped to the column's classes [0, nh)), j = 0 .. ncls;
// bnd[0] = start, bnd[j] = end for j >= ncls.
struct WinRange {
    uint32_t bnd[kMaxClasses + 1];
    __device__ uint32_t start() const { return bnd[0]; }
};
__device__ __forceinline__ WinRange window_range(const uint32_t* __restrict__ boff, const WinGeom& wg, int w, uint32_t& end) {
    const int ls = w / wg.nwin, r = w - ls * wg.nwin;
    const int wa = r / wg.nw, wb = r - wa * wg.nw;
    const int base = (ls * wg.nw + wb) * wg.nh;
    const int h0 = (wa << (4 - wg.cls_log2)) - (8 >> wg.cls_log2);
    WinRange x;
#pragma unroll
    for (int j = 0; j <= kMaxClasses; ++j)
        x.bnd[j] = boff[base + min(max(h0 + min(j, wg.ncls), 0), wg.nh)];
    end = x.bnd[wg.ncls];
    return x;
}

// Work items.  A window with at most coop_min records is one wavefront item.  A larger one is
// taken by whole workgroups: pieces of at most coop_cap records whose chunks the workgroup's
// wavefronts share, their accumulators summed in LDS.  A window with several pieces writes one
// partial window per piece, and k_win_reduce adds them in piece order (so the result does not
// depend on timing).  Both queues run largest first, by octave classes: class c holds items
// with count * 2^c <= cap < count * 2^(c+1) (cap = coop_min for wavefront items, coop_cap for
// workgroup items; the last class also the smaller and empty ones), so that the wavefronts and
// workgroups finish together.  The default limits were measured (SKA-Low snapshot at 1024^2,
// whose 9216 windows hold 1 to 1.8 M records each): below ~1000 records a workgroup item spends
// a large part of its time summing the wavefronts' accumulators, and pieces much above ~25000
// records leave the last ones running alone at the end of the launch; a context's tuning
// (svdi_autotune) may replace them per grid size.
// icount: [kItemClasses][nwin_total] wavefront items, [kItemClasses][nwin_total] workgroup items,
// [nwin_total] partial slots; one exclusive scan gives the queue positions and slot numbers.
// totals[0] = records, [1] = wavefront items, [2] = (unused), [3] = workgroup items,
// [4] = max |v| bits, [5] = workgroup queue head, [6] = windows with several pieces.
constexpr int kItemClasses = 8;
constexpr uint32_t kCoopMinDefault = 1024;
constexpr uint32_t kCoopCapDefault = 24576;

__device__ __forceinline__ int item_class(uint32_t count, uint32_t cap) {
    if (count == 0) return kItemClasses - 1;
    int c = 0;
    while (c < kItemClasses - 1 && (static_cast<uint64_t>(count) << (c + 1)) <= cap) ++c;
    return c;
}

__device__ __forceinline__ uint32_t coop_pieces(const WinGeom& wg, uint32_t c) {
    return (c + wg.coop_cap - 1) / wg.coop_cap;
}

__global__ void __launch_bounds__(256)
k_win_item_counts(const uint32_t* __restrict__ boff, WinGeom wg, uint32_t* __restrict__ icount,
                  uint32_t* __restrict__ multi, uint32_t* __restrict__ totals) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    const int nwt = wg.nwin_total;
    if (w == 0) totals[0] = boff[wg.nbins];
    if (w >= nwt) return;
    uint32_t end;
    const WinRange x = window_range(boff, wg, w, end);
    const uint32_t c = end - x.start();
    for (int k = 0; k <= 2 * kItemClasses; ++k) icount[k * nwt + w] = 0u;
    if (c <= wg.coop_min) {
        icount[item_class(c, wg.coop_min) * nwt + w] = 1u;
        return;
    }
    const uint32_t full = c / wg.coop_cap, rem = c - full * wg.coop_cap;
    uint32_t* gcount = icount + kItemClasses * nwt;
    gcount[w] += full;
    if (rem > 0) gcount[item_class(rem, wg.coop_cap) * nwt + w] += 1u;
    const uint32_t pieces = coop_pieces(wg, c);
    if (pieces > 1) {
        icount[2 * kItemClasses * nwt + w] = pieces;
        multi[atomicAdd(&totals[6], 1u)] = static_cast<uint32_t>(w);
    }
}

// One thread per item: its (queue, class, window) entry of the counts is found by binary search
// over the offsets.  Workgroup piece q of class 0 below the window's full count is full piece q;
// any other is the remainder (piece `full`).
__global__ void __launch_bounds__(256)
k_win_item_write(const uint32_t* __restrict__ boff, const uint32_t* __restrict__ ioff,
                 WinGeom wg, WItem* __restrict__ items, uint32_t max_items, uint32_t* __restrict__ totals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const int nwt = wg.nwin_total;
    const int n = 2 * kItemClasses * nwt;
    const uint32_t nwave = ioff[kItemClasses * nwt];
    const uint32_t nitems = min(ioff[n], max_items);
    if (i == 0) {
        totals[1] = min(nwave, nitems);
        totals[3] = nitems - min(nwave, nitems);
    }
    if (i >= nitems) return;
    int lo = 0, hi = n;                         // largest j with ioff[j] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (ioff[mid] <= i) lo = mid;
        else hi = mid;
    }
    const bool coop = lo >= kItemClasses * nwt;
    const int j = coop ? lo - kItemClasses * nwt : lo;
    const int cls = j / nwt;
    const int w = j - cls * nwt;
    uint32_t end;
    const WinRange x = window_range(boff, wg, w, end);
    const uint32_t c = end - x.start();
    WItem t;
    t.win = w;
#pragma unroll
    for (int j = 0; j < kMaxClasses; ++j) t.bnd[j] = static_cast<int>(x.bnd[j]);
    t.start = static_cast<int>(x.start());
    t.count = static_cast<int>(c);
    t.slot = -1;
    if (coop) {
        const uint32_t full = c / wg.coop_cap;
        const uint32_t q = i - ioff[lo];
        const uint32_t piece = (cls == 0 && q < full) ? q : full;
        t.start = static_cast<int>(x.start() + piece * wg.coop_cap);
        t.count = static_cast<int>(piece < full ? wg.coop_cap : c - full * wg.coop_cap);
        if (coop_pieces(wg, c) > 1) t.slot = static_cast<int>(ioff[n + w] - ioff[n] + piece);
    }
    items[i] = t;
}

// Windows with several pieces: the sum of their partial windows, in piece order, into the
// planes.  One unit = (listed window, term) per workgroup of kReduceGroups x 256 threads: group g
// sums the pieces g, g + kReduceGroups, ... of every cell (the densest windows have over a
// hundred pieces, which one thread per cell would request one after another), and the groups'
// sums are added in group order, so the result does not depend on timing.
constexpr int kReduceGroups = 4;
__global__ void __launch_bounds__(kReduceGroups * 256)
k_win_reduce(const uint32_t* __restrict__ boff, const uint32_t* __restrict__ ioff,
             const uint32_t* __restrict__ multi, const uint32_t* __restrict__ totals, WinGeom wg,
             int G, int K, uint32_t max_slots, const float2* __restrict__ partial, float2* __restrict__ planes) {
    __shared__ float2 s_sum[kReduceGroups - 1][256];
    const int n = 2 * kItemClasses * wg.nwin_total;
    const uint32_t units = totals[6] * static_cast<uint32_t>(K);
    const size_t plane_size = static_cast<size_t>(G) * G;
    const size_t piece_size = static_cast<size_t>(K) * kWin * kWin;
    const int cell = threadIdx.x & 255, grp = threadIdx.x >> 8;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const int w = static_cast<int>(multi[u / K]);
        const int k = static_cast<int>(u % K);
        uint32_t end;
        const WinRange x = window_range(boff, wg, w, end);
        const uint32_t pieces = coop_pieces(wg, end - x.start());
        const uint32_t first = ioff[n + w] - ioff[n];
        if (first + pieces > max_slots) continue;          // never: the slot count bounds the pieces
        const float2* src = partial + (first + grp) * piece_size + static_cast<size_t>(k) * kWin * kWin + cell;
        const size_t step = kReduceGroups * piece_size;
        // This group's pieces, eight requested at a time, summed into four accumulators in a
        // fixed order.
        const uint32_t mine = pieces > static_cast<uint32_t>(grp) ? (pieces - grp + kReduceGroups - 1) / kReduceGroups : 0;
        float2 acc[4] = {};
        uint32_t p = 0;
        for (; p + 8 <= mine; p += 8) {
            float2 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = src[(p + j) * step];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                acc[j & 3].x += v[j].x;
                acc[j & 3].y += v[j].y;
            }
        }
        for (; p < mine; ++p) {
            const float2 v = src[p * step];
            acc[p & 3].x += v.x;
            acc[p & 3].y += v.y;
        }
        const float2 sum = make_float2((acc[0].x + acc[1].x) + (acc[2].x + acc[3].x),
                                       (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y));
        if (grp > 0) s_sum[grp - 1][cell] = sum;
        __syncthreads();
        if (grp == 0) {
            float2 t = sum;
#pragma unroll
            for (int g = 0; g < kReduceGroups - 1; ++g) {
                t.x += s_sum[g][cell].x;
                t.y += s_sum[g][cell].y;
            }
            const int ls = w / wg.nwin, r = w - ls * wg.nwin;
            const int wa = r / wg.nw, wb = r - wa * wg.nw;
            const int a = wa * kWin + cell / kWin, b = wb * kWin + cell % kWin;
            if (a < G && b < G)
                planes[(static_cast<size_t>(ls) * K + k) * plane_size + static_cast<size_t>(a) * G + b] = t;
        }
        __syncthreads();
    }
}

// (f16(a - h.x), f16(b - h.y)) with each difference formed exactly in FP32: one mixed-precision
// fma per value (v_fma_mixlo / mixhi: source 0 the f16 half of h, negated, times 1.0, plus the
// FP32 value).
__device__ __forceinline__ half2v residual_f16x2(half2v h, float a, float b) {
    half2v r;
    asm volatile("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(a));
    asm volatile("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(r) : "v"(h), "v"(b));
    return r;
}

// x = hi + lo: hi rounded to nearest by a packed conversion (two values per instruction), lo =
// f16(x - hi) by mixed-precision fmas: |x - hi - lo| <= 2^-12 |lo| <= 2^-23 |x| (x in f16's
// normal range).
__device__ __forceinline__ void split_hilo(const float4& x0, const float4& x1, half8& hi, half8& lo) {
    const float v[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
        const half2v h = __builtin_convertvector((float2v){v[i], v[i + 1]}, half2v);
        const half2v l = residual_f16x2(h, v[i], v[i + 1]);
        hi[i] = h[0];
        hi[i + 1] = h[1];
        lo[i] = l[0];
        lo[i + 1] = l[1];
    }
}

// x = hi + lo in bf16: hi = RN(x), lo = RN(x - hi) (the difference exact in FP32), so
// |x - hi - lo| <= 2^-9 |x - hi| <= 2^-18 |x|; two values per packed conversion.
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
        const bf16x2 h = __builtin_convertvector((float2v){x[i], x[i + 1]}, bf16x2);
        const bf16x2 l = __builtin_convertvector((float2v){x[i] - static_cast<float>(h[0]),
                                                           x[i + 1] - static_cast<float>(h[1])}, bf16x2);
        hi[i] = h[0];
        hi[i + 1] = h[1];
        lo[i] = l[0];
        lo[i + 1] = l[1];
    }
}

// One window row a with bf16 operands: B_a = ku_a kv formed in FP32 and split into bf16 hi + lo,
// then D_a += C_hi B_hi + C_lo B_hi + C_hi B_lo.
__device__ __forceinline__ void window_row_bf16(floatx4& d, const float* __restrict__ u_row, const float (&vf)[8],
                                                const bf16x8& a_hi, const bf16x8& a_lo) {
    const float4 u0 = *reinterpret_cast<const float4*>(u_row), u1 = *reinterpret_cast<const float4*>(u_row + 4);
    const float p[8] = {u0.x * vf[0], u0.y * vf[1], u0.z * vf[2], u0.w * vf[3],
                        u1.x * vf[4], u1.y * vf[5], u1.z * vf[6], u1.w * vf[7]};
    bf16x8 b_hi, b_lo;
    split_bf16(p, b_hi, b_lo);
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, b_hi, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo, b_hi, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi, b_lo, d, 0, 0, 0);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One window row a.  kPrecSplit3: B_a[r][b] = ku_r[a] kv_r[b] from the factors split into
// f16 hi + lo (ku = uh + ul, kv = vh + vl), in packed f16: P = f16(uh vh), E = uh vh - P exactly
// (an fma of 11-bit significands), lo = f16(uh vl + ul vh + E), hi = P, so |ku kv - hi - lo| <=
// 2^-11 |lo| + |ul vl| ~ 2^-22 |ku kv| (products below 2^-24 / 2^-22 of the largest lose bits to
// f16 subnormals: contributions under 1e-12 of the peak); then D_a += C B_a as hi*hi + lo*hi +
// hi*lo.  Otherwise B_a = f16(uh vh) (relative error <= 3 x 2^-11 against ku kv, from the three
// roundings) and D_a += C_hi B_a [+ C_lo B_a].
template <int P>
__device__ __forceinline__ void window_row_regs(floatx4& d, const half8& uh, const half8& ul, const half8& vh,
                                                const half8& vl, const half8& a_hi, const half8& a_lo) {
    if constexpr (P == kPrecSplit3) {
        const half8 p = uh * vh;
        const half8 e = __builtin_elementwise_fma(uh, vh, -p);
        const half8 b_lo = __builtin_elementwise_fma(ul, vh, __builtin_elementwise_fma(uh, vl, e));
        const half8 b_hi = p;
        d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, b_hi, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_lo, b_hi, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, b_lo, d, 0, 0, 0);
    } else {
        const half8 b = uh * vh;
        d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, b, d, 0, 0, 0);
        if constexpr (P == kPrecSplitC) d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_lo, b, d, 0, 0, 0);
    }
}

// Per-wavefront constants of the window kernel.
struct WinLane {
    _Float16* s_kuh;    // ku hi, lo: [32 rows a = -8 .. 23][32 samples], f16 (lo: kPrecSplit3 only)
    _Float16* s_kul;
    _Float16* s_kvh;    // kv hi, lo: [32 columns b = -8 .. 23][32 samples], f16
    _Float16* s_kvl;
    float* s_c;         // [16 m][32 samples]
    _Float16* s_kh;     // this lane's axis (ku or kv): hi and lo planes
    _Float16* s_kl;
    float* s_kuf;       // FP32 ES planes (kPrecBf16): ku, kv, and this lane's axis
    float* s_kvf;
    float* s_kf;
    int lane, rl, axis, g8, n16;
    int vexp;           // log2 of the c scale
    float inv;
};

// One chunk: the records [c0, c0 + 32) of item `it` (this lane's record in q) contracted
// into the accumulators d for the terms kg .. kg + 7.
template <int P>
__device__ __forceinline__ void window_chunk(const WinLane& L, const WItem& it, int c0, int kg, int K,
                                             int cls_log2, const uint4& q, floatx4 (&d)[kWin]) {
    using T = WinTraits<P>;
    const bool valid = c0 + L.rl < it.count;
    // Classes of the chunk's first and last records (wave-uniform), and the 4-row quarters of
    // the window that they touch: class j covers rows [c d, c d + c + 7], d = j - 8 / c.
    const int i0 = it.start + c0, i1 = min(it.start + it.count, i0 + kChunk);
    // (i >= bnd as the sign of bnd - 1 - i, so that the sums stay in scalar registers; bnd[j] for
    // j >= ncls is the window's end, which no record reaches.)
    int j_first = 0, j_last = 0;
#pragma unroll
    for (int j = 1; j < kMaxClasses; ++j) {
        j_first += static_cast<int>((static_cast<uint32_t>(it.bnd[j]) + ~static_cast<uint32_t>(i0)) >> 31);
        j_last += static_cast<int>((static_cast<uint32_t>(it.bnd[j]) - static_cast<uint32_t>(i1)) >> 31);
    }
    const int cls = 1 << cls_log2, dlo = -(8 >> cls_log2);
    const int q_lo = max(0, ((j_first + dlo) * cls) >> 2);
    const int q_hi = min(3, ((j_last + dlo) * cls + cls + 7) >> 2);
    // ES factors along this lane's axis, as f16 (hi + lo for kPrecSplit3): the axis' 32 lanes
    // zero the window's 16 rows / columns of the planes (a 64-byte row per lane, or half a row
    // with one plane), then every lane writes its sample's 9 cells ceil(p - 4) .. ceil(p - 4) + 8
    // (padded rows take those outside the window; the ninth is nonzero only for an integer p).
    {
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (T::kBf16) {
            float* zrow = L.s_kf - L.rl + (kPad + (L.rl & 15)) * kStrideF + ((L.rl & 16) ? 16 : 0);
#pragma unroll
            for (int k = 0; k < 4; ++k) reinterpret_cast<uint4*>(zrow)[k] = z;
        } else if constexpr (T::kEsLo) {
            _Float16* zrow = ((L.rl & 16) ? L.s_kl : L.s_kh) - L.rl + (kPad + (L.rl & 15)) * kStrideH;
#pragma unroll
            for (int k = 0; k < 4; ++k) reinterpret_cast<uint4*>(zrow)[k] = z;
        } else {
            _Float16* zrow = L.s_kh - L.rl + (kPad + (L.rl & 15)) * kStrideH + ((L.rl & 16) ? 16 : 0);
#pragma unroll
            for (int k = 0; k < 2; ++k) reinterpret_cast<uint4*>(zrow)[k] = z;
        }
    }
    // Single-f16 B: each sample's two ES factors are scaled by its own factors du, dv in [1, 2)
    // (from the visibility and s bits of its record) and c by 1 / (du dv).  Rounding ku, kv and
    // their product to f16 then errs by an amount that is independent from sample to sample and
    // averages out over the samples of a cell; with a fixed scale the kernel's flat peak sits at
    // one f16 value for every sample and the rounding is biased (about 4e-6 of the gridded sum).
    float dith_u = 1.0f, dith_v = 1.0f;
    if constexpr (T::kDither) {
        const uint32_t h = (q.x ^ q.y ^ q.z) * 0x9E3779B1u;
        dith_u = 1.0f + static_cast<float>(h >> 24) * (1.0f / 256.0f);
        dith_v = 1.0f + static_cast<float>((h >> 16) & 0xffu) * (1.0f / 256.0f);
    }
    // A record's row position is relative to its class's home window row (the first window row
    // that reads the class); this window reads the classes j >= 16 / c as the row above their
    // home, 16 rows up.
    const int foreign = cls_log2 == 2 ? it.bnd[4] : it.bnd[2];
    if (valid) {
        const int i = i0 + L.rl;
        const uint32_t pw = L.axis ? q.w : q.z;
        const int fl = rec_cell(pw) + (L.axis == 0 && i >= foreign ? kWin : 0);
        const uint32_t fb = rec_frac_bits(pw);
        const int first = min(max(fl, -kPad), kWin - 1);
        // Cell j is at x = (F + j) / 4 - 1 (F the record's fraction in [0, 1)): for j < 8 the
        // quarter-support polynomial pair (j, 7 - j) in t = F - 1/2, for j = 8 ES(1) when F = 0
        // and zero otherwise.
        const float t = fmaf(static_cast<float>(fb), kRecFracInv, -0.5f), t2 = t * t;
        const float dith = L.axis ? dith_v : dith_u;
        float v[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float e = kEsPolyEven[j][kEsPolyDegree / 2];
#pragma unroll
            for (int i = kEsPolyDegree / 2 - 1; i >= 0; --i) e = fmaf(e, t2, kEsPolyEven[j][i]);
            float o = kEsPolyOdd[j][kEsPolyDegree / 2 - 1];
#pragma unroll
            for (int i = kEsPolyDegree / 2 - 2; i >= 0; --i) o = fmaf(o, t2, kEsPolyOdd[j][i]);
            if constexpr (T::kDither) {
                e *= dith;
                o *= dith;
            }
            v[j] = fmaf(t, o, e);
            v[7 - j] = fmaf(-t, o, e);
        }
        if constexpr (T::kBf16) {
            float* df = L.s_kf + (first + kPad) * kStrideF;
#pragma unroll
            for (int j = 0; j < 8; ++j) df[j * kStrideF] = v[j];
            df[8 * kStrideF] = fb == 0u ? kEsPolyEdge : 0.0f;
        } else {
        _Float16* dh = L.s_kh + (first + kPad) * kStrideH;
        // hi of the pair (j, 7 - j) by one packed round-to-nearest conversion, each lo = f16(v - hi)
        // with the difference formed in FP32 (one mixed-precision fma per value).
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const half2v h = __builtin_convertvector((float2v){v[j], v[7 - j]}, half2v);
            dh[j * kStrideH] = h[0];
            dh[(7 - j) * kStrideH] = h[1];
            if constexpr (T::kEsLo) {
                _Float16* dl = L.s_kl + (first + kPad) * kStrideH;
                const half2v l = residual_f16x2(h, v[j], v[7 - j]);
                dl[j * kStrideH] = l[0];
                dl[(7 - j) * kStrideH] = l[1];
            }
        }
        const bool edge = fb == 0u;
        if constexpr (T::kEsLo) dh[8 * kStrideH] = edge ? kEsPolyEdgeHi : static_cast<_Float16>(0.0f);
        else dh[8 * kStrideH] = static_cast<_Float16>(edge ? kEsPolyEdge * dith : 0.0f);
        if constexpr (T::kEsLo) {
            _Float16* dl = L.s_kl + (first + kPad) * kStrideH;
            dl[8 * kStrideH] = edge ? kEsPolyEdgeLo : static_cast<_Float16>(0.0f);
        }
        }
    }
    // c rows 2k + axis for the terms kg .. kg + 7.  Terms k >= K (when K is not a multiple of 8)
    // are contracted like the others and dropped at write-out (T_k stays in [-1, 1]).
    {
        float vs = valid ? rec_v(L.axis ? q.y : q.x, q.w, L.vexp) : 0.0f;
        if constexpr (T::kDither) vs *= __builtin_amdgcn_rcpf(dith_u * dith_v);
        const float s = rec_s(q.x, q.y, q.z), s2 = 2.0f * s;
        float t0 = 1.0f, t1 = s;                      // T_k, T_{k+1}
        for (int k = 0; k < kg; ++k) {
            const float t2 = s2 * t1 - t0;
            t0 = t1;
            t1 = t2;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            L.s_c[(2 * j + L.axis) * kStride + L.rl] = t0 * vs;
            const float t2 = s2 * t1 - t0;
            t0 = t1;
            t1 = t2;
        }
    }
    wave_lds_sync();

    // A (16 x 32): lane holds A[n16][g8 + i]; B (32 x 16): lane holds B[g8 + i][n16];
    // D (16 x 16): lane holds D[4 (lane / 16) + i][n16].
    if constexpr (T::kBf16) {
        bf16x8 a_hi, a_lo;
        {
            const float* row = L.s_c + L.n16 * kStride + L.g8;
            const float4 x0 = *reinterpret_cast<const float4*>(row), x1 = *reinterpret_cast<const float4*>(row + 4);
            const float x[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            split_bf16(x, a_hi, a_lo);
        }
        float vf[8];
        {
            const float* vrow = L.s_kvf + (kPad + L.n16) * kStrideF + L.g8;
            const float4 v0 = *reinterpret_cast<const float4*>(vrow), v1 = *reinterpret_cast<const float4*>(vrow + 4);
            vf[0] = v0.x; vf[1] = v0.y; vf[2] = v0.z; vf[3] = v0.w;
            vf[4] = v1.x; vf[5] = v1.y; vf[6] = v1.z; vf[7] = v1.w;
        }
        const float* ub = L.s_kuf + kPad * kStrideF + L.g8;
#pragma unroll
        for (int qd = 0; qd < 4; ++qd) {
            if (qd >= q_lo && qd <= q_hi) {
#pragma unroll
                for (int a = 4 * qd; a < 4 * qd + 4; ++a) window_row_bf16(d[a], ub + a * kStrideF, vf, a_hi, a_lo);
            }
        }
        wave_lds_sync();
        return;
    }
    half8 a_hi, a_lo;
    {
        const float* row = L.s_c + L.n16 * kStride + L.g8;
        const float4 x0 = *reinterpret_cast<const float4*>(row), x1 = *reinterpret_cast<const float4*>(row + 4);
        if constexpr (T::kCLo) {
            split_hilo(x0, x1, a_hi, a_lo);
        } else {
            a_hi = __builtin_convertvector((floatx8){x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w}, half8);
            a_lo = a_hi;
        }
    }
    const half8 vh = *reinterpret_cast<const half8*>(L.s_kvh + (kPad + L.n16) * kStrideH + L.g8);
    half8 vl = vh;
    if constexpr (T::kEsLo) vl = *reinterpret_cast<const half8*>(L.s_kvl + (kPad + L.n16) * kStrideH + L.g8);
    const _Float16* uhb = L.s_kuh + kPad * kStrideH + L.g8;
    const _Float16* ulb = T::kEsLo ? L.s_kul + kPad * kStrideH + L.g8 : uhb;
#pragma unroll
    for (int qd = 0; qd < 4; ++qd) {
        if (qd >= q_lo && qd <= q_hi) {
            // Each row's ES values are requested one row ahead, so that their LDS latency
            // overlaps the previous row's products and MFMAs.
            half8 uh_n = *reinterpret_cast<const half8*>(uhb + 4 * qd * kStrideH);
            half8 ul_n = uh_n;
            if constexpr (T::kEsLo) ul_n = *reinterpret_cast<const half8*>(ulb + 4 * qd * kStrideH);
#pragma unroll
            for (int a = 4 * qd; a < 4 * qd + 4; ++a) {
                const half8 uh = uh_n, ul = ul_n;
                if (a + 1 < 4 * qd + 4) {
                    uh_n = *reinterpret_cast<const half8*>(uhb + (a + 1) * kStrideH);
                    if constexpr (T::kEsLo) ul_n = *reinterpret_cast<const half8*>(ulb + (a + 1) * kStrideH);
                }
                window_row_regs<P>(d[a], uh, ul, vh, vl, a_hi, a_lo);
            }
        }
    }
    wave_lds_sync();
}

// Destination of the window values of (row a, term k, column n16): the planes for a whole
// window, the partial slot ([k][16][16]) for a piece; out-of-grid cells and k >= K go to the
// scratch line `trash` so that every lane issues the same stores.
struct WinOut {
    float2* base;
    int kstride, astride;
    bool piece;
};
__device__ __forceinline__ WinOut window_out(const WItem& it, const WinGeom& wg, int K, int G, float2* planes,
                                             float2* partial, int n16) {
    const int ls = it.win / wg.nwin, w = it.win - ls * wg.nwin;
    const int wa = w / wg.nw, wb = w - wa * wg.nw;
    const size_t plane_size = static_cast<size_t>(G) * G;
    WinOut o;
    o.piece = it.slot >= 0;
    o.base = o.piece ? partial + static_cast<size_t>(it.slot) * K * kWin * kWin + n16
                     : planes + static_cast<size_t>(ls) * K * plane_size + static_cast<size_t>(wa) * kWin * G + wb * kWin + n16;
    // Offsets from base fit in 32 bits: k * G^2 + 15 G < 2^31 for K <= 32, G <= 8192.
    o.kstride = o.piece ? kWin * kWin : static_cast<int>(plane_size);
    o.astride = o.piece ? kWin : G;
    return o;
}

__device__ __forceinline__ void window_store(const WinOut& o, const WItem& it, const WinGeom& wg, int G, int K,
                                             int a, int k, int n16, float inv, float2 v, float2* trash, int lane) {
    const int w = it.win % wg.nwin;
    const int ag = (w / wg.nw) * kWin + a, bg = (w % wg.nw) * kWin + n16;
    const float sg = ((ag + bg) & 1) ? -inv : inv;
    const bool ok = k < K && (o.piece || (ag < G && bg < G));
    float2* dst = ok ? o.base + (k * o.kstride + a * o.astride) : trash + lane;
    *dst = make_float2(v.x * sg, v.y * sg);
}

// One job of the window kernel: one slab batch of one snapshot, binned (records, item lists and
// totals in its call's workspace), with its own planes.  A launch takes up to kMaxJobs of them.
constexpr int kMaxJobs = kMaxWindowJobs;
struct WinJob {
    const WItem* items;
    uint32_t* totals;
    WRecords rec;
    float2* planes;
    float2* partial;
    float2* trash;
    WinGeom wg;
    int K, G;
};
struct WinJobs {
    WinJob job[kMaxJobs];
    int count;
};

// The c scale of a job from the largest |v| of its call: |c| <= 2^14 after scaling.
__device__ __forceinline__ void job_scale(const WinJob& J, WinLane& L) {
    const float vmax = __uint_as_float(J.totals[4]);
    int e = 0;
    if (vmax > 0.0f && vmax <= 3.0e38f) (void)frexpf(vmax, &e);
    e = min(max(e, -100), 100);
    L.vexp = 14 - e;
    L.inv = ldexpf(1.0f, e - 14 - kEsScaleLog2Sq);
}

// The job of global item index g of a queue whose per-job starts are off[0 .. count] (LDS):
// the largest j with off[j] <= g (wave-uniform).
__device__ __forceinline__ int job_of(const int* off, int count, int g) {
    int j = 0;
    while (j + 1 < count && __builtin_amdgcn_readfirstlane(off[j + 1]) <= g) ++j;
    return j;
}

// Persistent workgroups, one per CU, over the items of all the launch's jobs.  First the
// workgroup items of every job, from one queue: all wavefronts of the workgroup claim the item's
// chunks from a shared counter, then sum their accumulators in LDS (over the staging area) and
// the workgroup writes the window once.  Then the wavefront items of every job, each wavefront
// on its own: the next item is claimed while the current one is contracted and its first records
// are requested before the current window is written, so the write-out overlaps the next item's
// loads (vector memory operations complete in order).  The terms are done eight at a time (one
// 16-row block of m = 2k + ri).
template <int P>
__global__ void __launch_bounds__(WinTraits<P>::kWaves * 64)
k_win_grid(WinJobs jobs) {
    using T = WinTraits<P>;
    constexpr int kWaves = T::kWaves;
    extern __shared__ __align__(16) float stage_all[];
    __shared__ int s_item, s_chunk, s_witem;
    __shared__ int s_coop_off[kMaxJobs + 1], s_wave_off[kMaxJobs + 1];
    WinLane L;
    const int wave = threadIdx.x >> 6;
    L.lane = threadIdx.x & 63;
    L.s_kuh = reinterpret_cast<_Float16*>(stage_all + wave * T::kStageFloats);
    L.s_kuf = L.s_kvf = L.s_kf = nullptr;
    if constexpr (T::kBf16) {
        L.s_kuf = stage_all + wave * T::kStageFloats;
        L.s_kvf = L.s_kuf + kPlaneFloats;
        L.s_c = L.s_kvf + kPlaneFloats;
        L.s_kul = L.s_kvh = L.s_kvl = L.s_kuh;
    } else if constexpr (T::kEsLo) {
        L.s_kul = L.s_kuh + kPlaneHalves;
        L.s_kvh = L.s_kul + kPlaneHalves;
        L.s_kvl = L.s_kvh + kPlaneHalves;
        L.s_c = reinterpret_cast<float*>(L.s_kvl + kPlaneHalves);
    } else {
        L.s_kvh = L.s_kuh + kPlaneHalves;
        L.s_kul = L.s_kuh;
        L.s_kvl = L.s_kvh;
        L.s_c = reinterpret_cast<float*>(L.s_kvh + kPlaneHalves);
    }
    L.rl = L.lane & 31;
    L.axis = L.lane >> 5;
    L.g8 = 8 * (L.lane >> 4);
    L.n16 = L.lane & 15;
    L.s_kh = (L.axis ? L.s_kvh : L.s_kuh) + L.rl;
    L.s_kl = (L.axis ? L.s_kvl : L.s_kul) + L.rl;
    if constexpr (T::kBf16) L.s_kf = (L.axis ? L.s_kvf : L.s_kuf) + L.rl;
    const int lane = L.lane;
    const int njobs = jobs.count;

    // The jobs' item counts as queue offsets (workgroup items first, then wavefront items, each
    // job's list in its largest-first order), and the first workgroup item claimed from the one
    // queue head (job 0's totals[5]).  The claimed index is read back into a scalar register, so
    // the loop exit is wave-uniform by construction and every wavefront runs the same barriers:
    // were it left in a vector register, the compiler could treat the exit as divergent and let
    // lanes other than thread 0 iterate without the claim.  Within an item the wavefronts claim
    // chunks from an LDS counter (the SIMDs issue their oldest wavefront first, so an even split
    // would leave the item waiting for the youngest); each claims its next chunk before
    // contracting the current one, and requests that chunk's records at once.
    const int wave_id = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    if (threadIdx.x == 0) {
        int c = 0, w = 0;
        for (int j = 0; j < njobs; ++j) {
            s_coop_off[j] = c;
            s_wave_off[j] = w;
            c += static_cast<int>(jobs.job[j].totals[3]);
            w += static_cast<int>(jobs.job[j].totals[1]);
        }
        s_coop_off[njobs] = c;
        s_wave_off[njobs] = w;
        s_chunk = 0;
        s_witem = 0;
        s_item = static_cast<int>(atomicAdd(&jobs.job[0].totals[5], 1u));
    }
    auto claim_chunk = [&]() -> int {
        int v = 0;
        if (lane == 0) v = atomicAdd(&s_chunk, kChunk);
        return __builtin_amdgcn_readfirstlane(v);
    };
    for (;;) {
        __syncthreads();
        const int gi = __builtin_amdgcn_readfirstlane(s_item);
        const int ncoop = __builtin_amdgcn_readfirstlane(s_coop_off[njobs]);
        __syncthreads();
        if (gi >= ncoop) break;
        // The next item is claimed now; wavefront 0 publishes it after the last barrier of this one.
        uint32_t gnext = 0;
        if (wave_id == 0 && lane == 0) gnext = atomicAdd(&jobs.job[0].totals[5], 1u);
        const int jj = job_of(s_coop_off, njobs, gi);
        const WinJob& J = jobs.job[jj];
        job_scale(J, L);
        const int K = J.K, G = J.G;
        const int nwave_j = __builtin_amdgcn_readfirstlane(s_wave_off[jj + 1] - s_wave_off[jj]);
        const WItem it = J.items[nwave_j + (gi - __builtin_amdgcn_readfirstlane(s_coop_off[jj]))];
        for (int kg = 0; kg < K; kg += 8) {
            floatx4 d[kWin];
#pragma unroll
            for (int a = 0; a < kWin; ++a) d[a] = floatx4{0.f, 0.f, 0.f, 0.f};
            int c0 = claim_chunk();
            uint4 q_next = make_uint4(0u, 0u, 0u, 0u);
            if (c0 < it.count) q_next = J.rec.get(it.start + min(c0 + L.rl, it.count - 1));
            while (c0 < it.count) {
                const uint4 q = q_next;
                const int c1 = claim_chunk();
                if (c1 < it.count) q_next = J.rec.get(it.start + min(c1 + L.rl, it.count - 1));
                window_chunk<P>(L, it, c0, kg, K, J.wg.cls_log2, q, d);
                c0 = c1;
            }
            // Sum the wavefronts' accumulators in LDS (over the staging area), in a fixed order:
            // wavefronts 0 .. S-1 store theirs in slots 0 .. S-1, then each following group of S
            // wavefronts adds theirs to slot w mod S, one group per round, and the write sums the
            // S slots.  Slot layout red4[slot][a][lane] (float4 = the lane's
            // D[4 (lane / 16) + i][n16], i = 0..3).  The chunk counter is reset for the next term
            // block or item once every wavefront is done with it.
            constexpr int kSlots = T::kSlots;
            float4* red4 = reinterpret_cast<float4*>(stage_all);
            __syncthreads();
            if (wave_id < kSlots) {
#pragma unroll
                for (int a = 0; a < kWin; ++a)
                    red4[(wave_id * kWin + a) * 64 + lane] = make_float4(d[a][0], d[a][1], d[a][2], d[a][3]);
            }
            if (threadIdx.x == 0) s_chunk = 0;
            __syncthreads();
#pragma unroll
            for (int round = 1; round < kWaves / kSlots; ++round) {
                if (wave_id >= round * kSlots && wave_id < (round + 1) * kSlots) {
#pragma unroll
                    for (int a = 0; a < kWin; ++a) {
                        float4& r = red4[((wave_id - round * kSlots) * kWin + a) * 64 + lane];
                        const float4 v = r;
                        r = make_float4(v.x + d[a][0], v.y + d[a][1], v.z + d[a][2], v.w + d[a][3]);
                    }
                }
                __syncthreads();
            }
            // Write: (row a, source lane sl) -> terms kg + 2 (sl / 16) + h, h = 0, 1, column sl % 16.
            for (int v = threadIdx.x; v < kWin * 64; v += kWaves * 64) {
                const int a = v >> 6, sl = v & 63;
                float4 acc = red4[a * 64 + sl];
#pragma unroll
                for (int j = 1; j < kSlots; ++j) {
                    const float4 x = red4[(j * kWin + a) * 64 + sl];
                    acc = make_float4(acc.x + x.x, acc.y + x.y, acc.z + x.z, acc.w + x.w);
                }
                const WinOut os = window_out(it, J.wg, K, G, J.planes, J.partial, sl & 15);
                window_store(os, it, J.wg, G, K, a, kg + 2 * (sl >> 4), sl & 15, L.inv, make_float2(acc.x, acc.y), J.trash, sl);
                window_store(os, it, J.wg, G, K, a, kg + 2 * (sl >> 4) + 1, sl & 15, L.inv, make_float2(acc.z, acc.w), J.trash, sl);
            }
            __syncthreads();
        }
        if (wave_id == 0 && lane == 0) s_item = static_cast<int>(gnext);
    }

    // Wavefront items of all jobs, one list.  Workgroup g takes items g, g + G, g + 2G, ... of it,
    // G = the number of workgroups, so the CUs get equal shares without a global queue; its
    // wavefronts claim them from an LDS counter, each claiming its next item while it contracts
    // the current one.
    const int nwave = __builtin_amdgcn_readfirstlane(s_wave_off[njobs]);
    const int nwg = static_cast<int>(gridDim.x);
    const int wg_items = nwave > static_cast<int>(blockIdx.x)
                             ? (nwave - static_cast<int>(blockIdx.x) + nwg - 1) / nwg : 0;
    auto claim_item = [&]() -> int {
        int v = 0;
        if (lane == 0) v = atomicAdd(&s_witem, 1);
        return v;
    };
    int ticket = claim_item();
    const int j0 = __builtin_amdgcn_readfirstlane(ticket);
    if (j0 >= wg_items) return;
    int item = static_cast<int>(blockIdx.x) + j0 * nwg;
    int jj = job_of(s_wave_off, njobs, item);
    WItem it = jobs.job[jj].items[item - __builtin_amdgcn_readfirstlane(s_wave_off[jj])];
    // Records of the next chunk, loaded while the current one is contracted.  Every lane loads
    // (lanes past the end re-read the last record and ignore it), so the loads land in place.
    uint4 q_next = jobs.job[jj].rec.get(it.count > 0 ? it.start + min(L.rl, it.count - 1) : 0);
    for (;;) {
        ticket = claim_item();
        const WinJob& J = jobs.job[jj];
        job_scale(J, L);
        const int K = J.K, G = J.G;
        int next = nwave, jn = jj;
        WItem it_next = it;
        const WinOut o = window_out(it, J.wg, K, G, J.planes, J.partial, L.n16);
        for (int kg = 0; kg < K; kg += 8) {
            floatx4 d[kWin];
#pragma unroll
            for (int a = 0; a < kWin; ++a) d[a] = floatx4{0.f, 0.f, 0.f, 0.f};
            for (int c0 = 0; c0 < it.count; c0 += kChunk) {
                const uint4 q = q_next;
                if (c0 + kChunk < it.count) q_next = J.rec.get(it.start + min(c0 + kChunk + L.rl, it.count - 1));
                window_chunk<P>(L, it, c0, kg, K, J.wg.cls_log2, q, d);
            }
            // What comes next: this item's next term block, or the next item (possibly of
            // another job).  Its first records are requested now, ahead of the stores below.
            {
                int pstart = it.start, pcount = it.count;
                const WRecords* prec = &J.rec;
                if (kg + 8 >= K) {
                    const int tn = __builtin_amdgcn_readfirstlane(ticket);
                    next = tn < wg_items ? static_cast<int>(blockIdx.x) + tn * nwg : nwave;
                    if (next < nwave) {
                        jn = job_of(s_wave_off, njobs, next);
                        it_next = jobs.job[jn].items[next - __builtin_amdgcn_readfirstlane(s_wave_off[jn])];
                        pstart = it_next.start;
                        pcount = it_next.count;
                        prec = &jobs.job[jn].rec;
                    } else {
                        pcount = 0;
                    }
                }
                q_next = prec->get(pcount > 0 ? pstart + min(L.rl, pcount - 1) : 0);
            }
            // Rows m = 4 (lane / 16) + i are (re, im) of terms kg + 2 (lane / 16) and
            // kg + 2 (lane / 16) + 1; column n16.
#pragma unroll
            for (int a = 0; a < kWin; ++a)
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    window_store(o, it, J.wg, G, K, a, kg + 2 * (lane >> 4) + h, L.n16, L.inv,
                                 make_float2(d[a][2 * h], d[a][2 * h + 1]), J.trash, lane);
        }
        if (next >= nwave) break;
        item = next;
        jj = jn;
        it = it_next;
    }
}

}  // namespace

namespace {
// Row-class height (log2) and bin count of a grid: classes of 4 cells when that histogram fits in
// LDS, else of 8 (half the bins, so half the LDS bands when even those do not fit).
int window_class_log2(const GridParams& gp, int nw) {
    const size_t cols = static_cast<size_t>(gp.nslabs) * nw;
    return cols * ((gp.G + 3) / 4) <= static_cast<size_t>(kLdsBins) ? 2 : 3;
}
size_t window_bins(const GridParams& gp, int nw, int cls_log2) {
    const size_t nh = (static_cast<size_t>(gp.G) + (1u << cls_log2) - 1) >> cls_log2;
    return static_cast<size_t>(gp.nslabs) * nw * nh;
}
}  // namespace

// The class height in use: the tuned one, else by LDS fit.
static int effective_class_log2(const GridParams& gp, int nw) {
    return gp.class_log2 == 2 || gp.class_log2 == 3 ? gp.class_log2 : window_class_log2(gp, nw);
}

bool grid_windows_supported(const SampleInput& in, const GridParams& gp) {
    if (gp.G > 8192 || gp.G < 1) return false;
    const int nw = (gp.G + kWin - 1) / kWin;
    if (window_bins(gp, nw, effective_class_log2(gp, nw)) >= (1ull << 30)) return false;
    return 2 * in.num_samples < (1ull << 31);             // records: at most two per sample
}

WindowJob bin_windows(svdi_context* ctx, const SampleInput& in, const GridParams& gp, float2* planes) {
    if (gp.K < 1 || gp.K > kMaxTerms) throw std::logic_error("term count out of range");
    if (!grid_windows_supported(in, gp)) throw UnsupportedError("grid or sample count too large for the window gridder");
    WinGeom wg;
    wg.nw = (gp.G + kWin - 1) / kWin;
    wg.cls_log2 = effective_class_log2(gp, wg.nw);
    wg.coop_min = gp.coop_min != 0 ? gp.coop_min : kCoopMinDefault;
    wg.coop_cap = gp.coop_cap != 0 ? gp.coop_cap : kCoopCapDefault;
    wg.ncls = 24 >> wg.cls_log2;
    wg.nh = (gp.G + (1 << wg.cls_log2) - 1) >> wg.cls_log2;
    wg.nwin = wg.nw * wg.nw;
    wg.nwin_total = wg.nwin * gp.nslabs;
    const size_t nbins_sz = window_bins(gp, wg.nw, wg.cls_log2);
    wg.nbins = static_cast<int>(nbins_sz);
    const size_t S = in.num_samples;
    const size_t max_records = 2 * S;                      // one row class, at most two window columns
    const size_t nbins = nbins_sz;
    const size_t nwt = static_cast<size_t>(wg.nwin_total);
    // Wavefront items: one per window at most.  Workgroup pieces: at most one per window plus
    // reads / coop_cap (a window reads each record at most twice).  Partial slots: only windows
    // with two or more pieces have them, at most 2 reads / coop_cap.
    const size_t max_reads = 2 * max_records;
    const size_t max_items = 2 * nwt + max_reads / wg.coop_cap + 1;
    const size_t max_slots = 2 * max_reads / wg.coop_cap + 1;
    unsigned blocks;
    if (in.from_snapshot) blocks = static_cast<unsigned>((in.num_rows + kBinRows - 1) / kBinRows);
    else blocks = static_cast<unsigned>((S + kBinThreads * kBinPerThread - 1) / (kBinThreads * kBinPerThread));
    // LDS bands of at most kLdsBins bins, unless the per-block histograms would pass
    // kMaxBlockHistBytes (then one global atomic per record).
    const size_t nbands = (nbins + kLdsBins - 1) / kLdsBins;
    wg.use_lds = static_cast<size_t>(blocks) * nbins * sizeof(uint32_t) <= kMaxBlockHistBytes ? 1 : 0;
    wg.band_bins = static_cast<int>((nbins + nbands - 1) / nbands);
    const unsigned bands = wg.use_lds ? static_cast<unsigned>(nbands) : 1u;
    const size_t nhist = wg.use_lds ? static_cast<size_t>(blocks) * nbins : 1;
    const size_t bin_lds = wg.use_lds ? static_cast<size_t>(wg.band_bins) * sizeof(uint32_t) : 0;
    const dim3 bin_grid(blocks, bands);

    // counts[0, nbins) | counts[nbins + 2, +8) = totals.
    uint32_t* counts = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinCounts, (nbins + 12) * sizeof(uint32_t)));
    // Per-block record offsets (the counts themselves when 32-bit), the (block, band) record
    // totals, then the 16-bit per-block counts when a block has fewer than 65536 samples (a
    // bin's count per block is at most its sample count).
    const size_t block_samples = in.from_snapshot ? static_cast<size_t>(kBinRows) * in.num_channels
                                                  : static_cast<size_t>(kBinThreads) * kBinPerThread;
    const bool count16 = wg.use_lds && block_samples < 65536;
    const size_t ntot = static_cast<size_t>(blocks) * bands;
    uint32_t* blockhist = static_cast<uint32_t*>(ctx->ws.get(
        Workspace::kRunOff, (nhist + ntot) * sizeof(uint32_t) + (count16 ? nhist * sizeof(uint16_t) : 0)));
    uint32_t* blocktot = blockhist + nhist;
    uint16_t* blockcount16 = count16 ? reinterpret_cast<uint16_t*>(blocktot + ntot) : nullptr;
    uint32_t* boff = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinOffsets, (nbins + 8) * sizeof(uint32_t)));
    uint32_t* cursor = static_cast<uint32_t*>(
        ctx->ws.get(Workspace::kBinCursor, (nbins + 8 + nwt + (wg.use_lds ? kBinSegs * nbins : 0)) * sizeof(uint32_t)));
    uint32_t* multi = cursor + nbins + 8;                  // windows with several pieces
    uint32_t* segsum = multi + nwt;                        // [kBinSegs][nbins]
    const size_t ncounts = (2 * kItemClasses + 1) * nwt;
    uint32_t* sums = static_cast<uint32_t*>(
        ctx->ws.get(Workspace::kScanSums, scan_blocks(std::max(nbins, ncounts)) * sizeof(uint32_t)));
    uint32_t* icount = static_cast<uint32_t*>(ctx->ws.get(Workspace::kItemCounts, (ncounts + 8) * sizeof(uint32_t)));
    uint32_t* ioff = static_cast<uint32_t*>(ctx->ws.get(Workspace::kItemOffsets, (ncounts + 8) * sizeof(uint32_t)));
    const size_t piece_floats2 = static_cast<size_t>(gp.K) * kWin * kWin;
    float2* partial = static_cast<float2*>(
        ctx->ws.get(Workspace::kWinPartial, (max_slots * piece_floats2 + 64) * sizeof(float2)));
    float2* trash = partial + max_slots * piece_floats2;
    uint32_t* totals = counts + nbins + 2;
    WItem* items = static_cast<WItem*>(ctx->ws.get(Workspace::kItems, max_items * sizeof(WItem)));
    WRecords rec;
    rec.r = static_cast<uint4*>(ctx->ws.get(Workspace::kRecPos, max_records * sizeof(uint4)));

    launch_fill(ctx, reinterpret_cast<float*>(counts), nbins + 12, 0.0f);
    // 16-bit LDS histograms (two bins per word) when there are also more blocks than CUs, so that
    // two can share a CU.  With one block per CU the halved LDS gains nothing and the packing
    // costs a little (snapshot rows at 64 channels: 35.9 against 33.9 us; preprocessed samples:
    // 50 against 64).
    const bool hist16 = count16 && blocks > static_cast<unsigned>(ctx->cu_count);
    const size_t count_lds = hist16 ? static_cast<size_t>((wg.band_bins + 1) / 2) * sizeof(uint32_t) : bin_lds;
    auto count = [&](auto kernel, auto* out) {
        allow_dynamic_lds(reinterpret_cast<const void*>(kernel), count_lds);
        kernel<<<bin_grid, kBinThreads, count_lds, ctx->stream>>>(in, gp, wg, counts, out, blocktot);
    };
    if (in.from_snapshot) {
        if (hist16) count(&k_win_count<true, true, uint16_t>, blockcount16);
        else if (count16) count(&k_win_count<true, false, uint16_t>, blockcount16);
        else count(&k_win_count<true, false, uint32_t>, blockhist);
    } else {
        if (hist16) count(&k_win_count<false, true, uint16_t>, blockcount16);
        else if (count16) count(&k_win_count<false, false, uint16_t>, blockcount16);
        else count(&k_win_count<false, false, uint32_t>, blockhist);
    }
    SVDI_LAUNCH_CHECK();
    const unsigned seg_grid = static_cast<unsigned>((kBinSegs * nbins + 255) / 256);
    if (wg.use_lds) {
        if (count16)
            k_bin_seg_sums<uint16_t><<<seg_grid, 256, 0, ctx->stream>>>(blockcount16, static_cast<int>(blocks),
                                                                       wg.nbins, segsum, counts);
        else
            k_bin_seg_sums<uint32_t><<<seg_grid, 256, 0, ctx->stream>>>(blockhist, static_cast<int>(blocks), wg.nbins,
                                                                       segsum, counts);
        SVDI_LAUNCH_CHECK();
    }
    launch_exclusive_scan(ctx, counts, boff, wg.use_lds ? nullptr : cursor, sums, nbins);
    if (wg.use_lds) {
        if (count16)
            k_bin_block_offsets<uint16_t><<<seg_grid, 256, 0, ctx->stream>>>(
                blockcount16, blockhist, static_cast<int>(blocks), wg.nbins, segsum, boff);
        else
            k_bin_block_offsets<uint32_t><<<seg_grid, 256, 0, ctx->stream>>>(
                blockhist, blockhist, static_cast<int>(blocks), wg.nbins, segsum, boff);
        SVDI_LAUNCH_CHECK();
    }
    k_win_item_counts<<<static_cast<unsigned>((nwt + 255) / 256), 256, 0, ctx->stream>>>(boff, wg, icount, multi, totals);
    SVDI_LAUNCH_CHECK();
    launch_exclusive_scan(ctx, icount, ioff, nullptr, sums, ncounts);
    k_win_item_write<<<static_cast<unsigned>((max_items + 255) / 256), 256, 0, ctx->stream>>>(
        boff, ioff, wg, items, static_cast<uint32_t>(max_items), totals);
    SVDI_LAUNCH_CHECK();
    const uint32_t limit = static_cast<uint32_t>(max_records);
    if (in.from_snapshot) {
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_win_scatter<true>), bin_lds);
        k_win_scatter<true><<<bin_grid, kBinThreads, bin_lds, ctx->stream>>>(in, gp, wg, blockhist, blocktot, cursor,
                                                                         rec, limit, totals + 4, ctx->dev_err);
    } else {
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_win_scatter<false>), bin_lds);
        k_win_scatter<false><<<bin_grid, kBinThreads, bin_lds, ctx->stream>>>(in, gp, wg, blockhist, blocktot, cursor,
                                                                          rec, limit, totals + 4, ctx->dev_err);
    }
    SVDI_LAUNCH_CHECK();

    WindowJob job;
    static_assert(sizeof(WinGeom) <= sizeof(job.geom), "window geometry does not fit the job");
    std::memcpy(job.geom, &wg, sizeof wg);
    job.items = items;
    job.totals = totals;
    job.records = rec.r;
    job.planes = planes;
    job.partial = partial;
    job.trash = trash;
    job.boff = boff;
    job.ioff = ioff;
    job.multi = multi;
    job.max_slots = static_cast<uint32_t>(max_slots);
    job.K = gp.K;
    job.G = gp.G;
    job.prec = gp.prec;
    return job;
}

void contract_windows(svdi_context* ctx, const WindowJob* jobs, int count) {
    if (count < 1 || count > kMaxJobs) throw std::logic_error("window job count out of range");
    WinJobs wj = {};
    wj.count = count;
    for (int j = 0; j < count; ++j) {
        if (jobs[j].prec != jobs[0].prec) throw std::logic_error("window jobs of one launch differ in precision");
        WinJob& d = wj.job[j];
        d.items = static_cast<const WItem*>(jobs[j].items);
        d.totals = jobs[j].totals;
        d.rec.r = static_cast<uint4*>(jobs[j].records);
        d.planes = jobs[j].planes;
        d.partial = jobs[j].partial;
        d.trash = jobs[j].trash;
        std::memcpy(&d.wg, jobs[j].geom, sizeof d.wg);
        d.K = jobs[j].K;
        d.G = jobs[j].G;
    }
    const size_t tok = stage_begin(ctx, 5);
    const unsigned grid = static_cast<unsigned>(std::max(1, ctx->cu_count - ctx->grid_reserve));
    auto launch = [&](auto prec) {
        constexpr int P = decltype(prec)::value;
        using T = WinTraits<P>;
        constexpr size_t lds = sizeof(float) * T::kStageFloats * T::kWaves;
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_win_grid<P>), lds);
        k_win_grid<P><<<grid, T::kWaves * 64, lds, ctx->stream>>>(wj);
    };
    const int prec = jobs[0].prec;
    if (prec == kPrecBf16) launch(std::integral_constant<int, kPrecBf16>());
    else if (prec == kPrecSingle) launch(std::integral_constant<int, kPrecSingle>());
    else if (prec == kPrecSplitC) launch(std::integral_constant<int, kPrecSplitC>());
    else launch(std::integral_constant<int, kPrecSplit3>());
    SVDI_LAUNCH_CHECK();
    stage_end(ctx, tok);
}

void reduce_windows(svdi_context* ctx, const WindowJob& job) {
    WinGeom wg;
    std::memcpy(&wg, job.geom, sizeof wg);
    k_win_reduce<<<2048, kReduceGroups * 256, 0, ctx->stream>>>(job.boff, job.ioff, job.multi, job.totals, wg, job.G,
                                                                job.K, job.max_slots, job.partial, job.planes);
    SVDI_LAUNCH_CHECK();
}

const uint32_t* launch_grid_windows(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                                    float2* planes) {
    const WindowJob job = bin_windows(ctx, in, gp, planes);
    contract_windows(ctx, &job, 1);
    reduce_windows(ctx, job);
    return job.totals;
}

}  // namespace svdi
