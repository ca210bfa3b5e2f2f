// Tiled ES gridder: samples binned into T x T grid tiles, each tile accumulated in LDS.
//
// Same sample model as the reference's fused_gridding (SVDIntegratedImager.cu:1218-1293), with
// a different schedule.  The reference adds every (cell, term) contribution to the G x G planes
// with a device atomic; here:
//   1. count:   each sample's clipped footprint is assigned to every tile it overlaps (a sample
//               straddling a tile edge is duplicated, so tiles own their cells exclusively) and
//               counted in a bucket (tile, exact footprint position within the tile);
//   2. scan:    exclusive scan of the bucket counts, so records are sorted by tile and, within
//               a tile, by footprint position; work items are contiguous ranges of one tile's
//               records of similar size (the dense uv-core tiles split into several items);
//   3. scatter: per-(sample, tile) records written to their bucket;
//   4. grid:    a persistent workgroup takes items from a queue; each run of same-position
//               records (at most 64, a "unit") is accumulated by one wavefront in registers
//               (lane = footprint cell, records broadcast with readlane) and added to the LDS
//               tile once per cell and term; the tile is then stored to the planes with its
//               checkerboard sign (split tiles add atomically into a zeroed tile instead).
// Tiles: T = 32 for K <= 8, T = 16 above (the K-plane LDS accumulator must fit twice per CU).
#include "sample.hpp"

namespace svdi {
namespace {

constexpr int kBinThreads = 256;
constexpr int kBinPerThread = 64;        // samples per thread, preprocessed input
constexpr int kMinItem = 4096;           // smallest cap on the records of one work item
constexpr int kMaxItem = 16384;          // largest cap
constexpr int kTileThreads = 512;
constexpr int kUnit = 64;                // records per accumulation unit (one wavefront)
constexpr int kScanPer = 4;              // elements per thread in the bucket scan
constexpr int kScanBlock = 1024 * kScanPer;

struct TileGeom {
    int T, ntx, ntiles, nbins, KS, nkeys;
    uint32_t nbuckets;
};

struct Item {
    int bin, start, count, split;
};

// Records, structure of arrays: q = (q1, q2, s, Re v), vy = Im v.  The footprint (position
// within the tile and extents) is the record's bucket key plus (na, nb) in `nanb`.
struct Records {
    float4* q;
    float* vy;
    uint8_t* nanb;           // na | nb << 4
};

__device__ __forceinline__ uint32_t bucket_of(const Footprint& f, const TileGeom& tg, int ta, int tb) {
    const int bin = f.local_slab * tg.ntiles + ta * tg.ntx + tb;
    const int key = (f.a0 - ta * tg.T + 8) * tg.KS + (f.b0 - tb * tg.T + 8);
    return static_cast<uint32_t>(bin) * tg.nkeys + key;
}

// Iterate the samples of this block.  Snapshot input: block = 256 consecutive rows, thread =
// row, every channel (positions of one baseline move radially with frequency, so a block
// touches few tiles).  Preprocessed input: 256 * kBinPerThread consecutive samples.
template <bool kSnap, class Fn>
__device__ __forceinline__ void for_block_samples(const SampleInput& in, Fn&& fn) {
    if constexpr (kSnap) {
        const size_t row = blockIdx.x * (size_t)kBinThreads + threadIdx.x;
        if (row >= in.num_rows) return;
        for (size_t chan = 0; chan < in.num_channels; ++chan) fn(row + chan * in.num_rows, row, chan);
    } else {
        const size_t base = blockIdx.x * (size_t)kBinThreads * kBinPerThread + threadIdx.x;
        for (int j = 0; j < kBinPerThread; ++j) {
            const size_t idx = base + static_cast<size_t>(j) * kBinThreads;
            if (idx >= in.num_samples) return;
            fn(idx, 0, 0);
        }
    }
}

// Every (sample, overlapped tile) pair of this block.  A clipped footprint starts in [0, G)
// and spans at most 9 <= T cells, so it overlaps at most 2 x 2 tiles, and its start relative to
// an overlapped tile lies in [-8, T).
template <bool kSnap, class Fn>
__device__ __forceinline__ void for_sample_tiles(const SampleInput& in, const GridParams& gp,
                                                 const TileGeom& tg, Fn&& fn) {
    for_block_samples<kSnap>(in, [&](size_t idx, size_t row, size_t chan) {
        Footprint f;
        if (!setup_sample<kSnap>(in, gp, idx, row, chan, f)) return;
        for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
            for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb) fn(f, ta, tb);
    });
}

__device__ __forceinline__ void write_record(const Records& rec, uint32_t slot, const Footprint& f) {
    rec.q[slot] = make_float4(f.q1, f.q2, f.s, f.vw.x);
    rec.vy[slot] = f.vw.y;
    rec.nanb[slot] = static_cast<uint8_t>(f.na | (f.nb << 4));
}

__device__ __forceinline__ bool same_footprint(const Footprint& a, const Footprint& b) {
    return a.a0 == b.a0 && a.b0 == b.b0 && a.na == b.na && a.nb == b.nb && a.local_slab == b.local_slab;
}

// Snapshot input, one thread per row: a short baseline keeps its footprint position over many
// consecutive channels, and the densest buckets (~1e5 records at the uv centre) are made of such
// runs.  A run of same-footprint samples costs one atomic per overlapped tile instead of one per
// sample.  `fn(run, count, chan_begin, chan_end)` is called once per run; the run's samples are
// the accepted ones among channels [chan_begin, chan_end).
template <class Fn>
__device__ __forceinline__ void for_row_runs(const SampleInput& in, const GridParams& gp, Fn&& fn) {
    const size_t row = blockIdx.x * (size_t)kBinThreads + threadIdx.x;
    if (row >= in.num_rows) return;
    Footprint run;
    uint32_t len = 0;
    size_t begin = 0;
    for (size_t chan = 0; chan < in.num_channels; ++chan) {
        Footprint f;
        if (!setup_sample<true>(in, gp, row + chan * in.num_rows, row, chan, f)) continue;
        if (len > 0 && same_footprint(f, run)) {
            ++len;
            continue;
        }
        if (len > 0) fn(run, len, begin, chan);
        run = f;
        len = 1;
        begin = chan;
    }
    if (len > 0) fn(run, len, begin, in.num_channels);
}

template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_bin_count(SampleInput in, GridParams gp, TileGeom tg, uint32_t* __restrict__ bcount) {
    if constexpr (kSnap) {
        for_row_runs(in, gp, [&](const Footprint& f, uint32_t len, size_t, size_t) {
            for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
                for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb)
                    atomicAdd(&bcount[bucket_of(f, tg, ta, tb)], len);
        });
    } else {
        for_sample_tiles<false>(in, gp, tg, [&](const Footprint& f, int ta, int tb) {
            atomicAdd(&bcount[bucket_of(f, tg, ta, tb)], 1u);
        });
    }
}

// Records are tile-independent (positions relative to the footprint start or absolute), so a
// sample overlapping several tiles writes the same record into each tile's bucket.
template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_bin_scatter(SampleInput in, GridParams gp, TileGeom tg, uint32_t* __restrict__ bcursor, Records rec) {
    if constexpr (kSnap) {
        const size_t row = blockIdx.x * (size_t)kBinThreads + threadIdx.x;
        for_row_runs(in, gp, [&](const Footprint& f, uint32_t len, size_t begin, size_t end) {
            uint32_t base[4];
            int nt = 0;
            for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
                for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb)
                    base[nt++] = atomicAdd(&bcursor[bucket_of(f, tg, ta, tb)], len);
            // Second sweep over the run's channels: the accepted ones are exactly its samples
            // (setup_sample is a pure function of the sample; the checks only bound the writes
            // to the reserved slots).
            uint32_t i = 0;
            for (size_t chan = begin; chan < end && i < len; ++chan) {
                Footprint g;
                if (!setup_sample<true>(in, gp, row + chan * in.num_rows, row, chan, g)) continue;
                if (!same_footprint(g, f)) continue;
                for (int t = 0; t < nt; ++t) write_record(rec, base[t] + i, g);
                ++i;
            }
        });
    } else {
        for_sample_tiles<false>(in, gp, tg, [&](const Footprint& f, int ta, int tb) {
            write_record(rec, atomicAdd(&bcursor[bucket_of(f, tg, ta, tb)], 1u), f);
        });
    }
}

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

// Exclusive scan of the bucket counts in three launches: per-block scans, a scan of the block
// totals, then the block offsets added (the scatter's cursors start as a copy).
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
    cursor[i] = v;
}

__device__ __forceinline__ uint32_t bin_records(const uint32_t* boff, const TileGeom& tg, int b) {
    return boff[static_cast<uint32_t>(b + 1) * tg.nkeys] - boff[static_cast<uint32_t>(b) * tg.nkeys];
}

// One block: work items over the bins.  An item holds at most `cap` records, cap = clamp(total
// / (4 x CUs), kMinItem, kMaxItem), so the dense uv-core tiles split into a few items of
// similar size (each adds its part of the tile atomically); every bin gets at least one item
// so every tile is written.  boff[nbuckets] holds the total.
// totals[0] = records, [1] = items, [2] = queue head of the tile kernel, [3] = cap.
__global__ void __launch_bounds__(1024)
k_make_items(const uint32_t* __restrict__ boff, TileGeom tg, Item* __restrict__ items, int max_items,
             int cu_count, uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_wave[16];
    const uint32_t total = boff[tg.nbuckets];
    const uint32_t fair = (total + 4u * cu_count - 1) / (4u * cu_count);
    const uint32_t cap = min(static_cast<uint32_t>(kMaxItem), max(static_cast<uint32_t>(kMinItem), fair));
    const int per = (tg.nbins + 1023) / 1024;
    const int lo = min(static_cast<int>(threadIdx.x) * per, tg.nbins), hi = min(lo + per, tg.nbins);
    uint32_t it = 0;
    for (int b = lo; b < hi; ++b) it += max(1u, (bin_records(boff, tg, b) + cap - 1) / cap);
    const uint32_t incl = block_scan_1024(it, s_wave);
    uint32_t base_it = incl - it;
    for (int b = lo; b < hi; ++b) {
        const uint32_t c = bin_records(boff, tg, b);
        const uint32_t start = boff[static_cast<uint32_t>(b) * tg.nkeys];
        const uint32_t n = max(1u, (c + cap - 1) / cap);
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t slot = base_it + q;
            if (slot < static_cast<uint32_t>(max_items)) {
                Item t;
                t.bin = b;
                t.start = static_cast<int>(start + q * cap);
                t.count = static_cast<int>(c > q * cap ? min(cap, c - q * cap) : 0u);
                t.split = n > 1;
                items[slot] = t;
            }
        }
        base_it += n;
    }
    if (threadIdx.x == 1023) {
        totals[0] = total;
        totals[1] = min(incl, static_cast<uint32_t>(max_items));
        totals[3] = cap;
    }
}

// Clear the tiles that are accumulated by several items (their items add atomically).
__global__ void __launch_bounds__(256)
k_zero_split(const uint32_t* __restrict__ boff, const uint32_t* __restrict__ totals, TileGeom tg,
             int G, int K, float2* __restrict__ planes) {
    const int bin = blockIdx.x;
    if (bin_records(boff, tg, bin) <= totals[3]) return;
    const int ls = bin / tg.ntiles, tile = bin - ls * tg.ntiles;
    const int ta = tile / tg.ntx, tb = tile - ta * tg.ntx;
    const size_t plane_size = static_cast<size_t>(G) * G;
    for (int e = threadIdx.x; e < K * tg.T * tg.T; e += 256) {
        const int k = e / (tg.T * tg.T), cc = e - k * tg.T * tg.T;
        const int a = ta * tg.T + cc / tg.T, b = tb * tg.T + cc % tg.T;
        if (a < G && b < G)
            planes[(static_cast<size_t>(ls) * K + k) * plane_size + static_cast<size_t>(a) * G + b] =
                make_float2(0.f, 0.f);
    }
}

// LDS column of cell (a, b) of a T x T accumulator row: T = 32 rows are XOR-swizzled so an
// 8 x 8 footprint hits 32 distinct banks per half-wavefront.
template <int T>
__device__ __forceinline__ int swz(int a, int b) {
    if constexpr (T == 32) return b ^ ((a & 3) << 3);
    else return b;
}

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kStageStride = 36;         // floats per staging row (32 records + 4 pad: no bank conflicts)
constexpr int kChunk = 32;               // records per MFMA reduction step (16x16x32 f16)
constexpr float kEsScale = 128.0f;       // per-axis ES scale: B = ES_a ES_b 2^14 stays in f16 range
constexpr int kEsScaleLog2Sq = 14;

// LDS of the tile kernels: the K planes (re, im) of one T x T tile, the unit list of the current
// item, and (MFMA gridder) a staging area per wavefront: ku[8], kv[8] (ES per axis) and
// c[2 KMAX] (Chebyshev-weighted visibility, row 2k + ri) for 32 records.
template <int T, int KMAX, int WAVES, bool kMfma>
struct TileLds {
    static constexpr int KS = T + 8;               // footprint starts in [-8, T) per axis
    static constexpr int NKEYS = KS * KS;
    static constexpr int MAXUNITS = NKEYS + kMaxItem / kUnit;
    static constexpr int kStageRows = 16 + 2 * KMAX;
    static constexpr size_t stage_floats = kMfma ? static_cast<size_t>(kStageRows) * kStageStride : 0;
    static constexpr size_t acc = 0;                                        // float [2][KMAX][T*T]
    static constexpr size_t units = acc + sizeof(float) * 2 * KMAX * T * T;  // u32 [MAXUNITS]
    static constexpr size_t stage = units + sizeof(uint32_t) * MAXUNITS;     // float [WAVES][rows][36]
    static constexpr size_t bytes = stage + sizeof(float) * stage_floats * WAVES;
};

// Wave 0 lists the units of an item: for each footprint key, the part of its bucket inside the
// item, cut into runs of at most 64 records.  Unit word: start within the item (14 bits) |
// count - 1 (7 bits) | key (11 bits).
template <int NKEYS>
__device__ __forceinline__ void build_units(const Item& it, const uint32_t* __restrict__ boff,
                                            uint32_t* units, uint32_t* s_nunits) {
    const int lane = threadIdx.x & 63;
    const uint32_t* bo = boff + static_cast<uint32_t>(it.bin) * NKEYS;
    const int i0 = it.start, i1 = it.start + it.count;
    uint32_t run_u = 0;
    for (int base = 0; base < NKEYS; base += 64) {
        const int k = base + lane;
        int lo = 0, c = 0;
        if (k < NKEYS) {
            lo = max(static_cast<int>(bo[k]), i0);
            c = max(min(static_cast<int>(bo[k + 1]), i1) - lo, 0);
        }
        const uint32_t u = (c + kUnit - 1) / kUnit;
        uint32_t su = u;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(su, d, 64);
            if (lane >= d) su += t;
        }
        const uint32_t uo = run_u + su - u;
        for (uint32_t q = 0; q < u; ++q)
            units[uo + q] = static_cast<uint32_t>(lo - i0 + static_cast<int>(q) * kUnit) |
                            (static_cast<uint32_t>(min(kUnit, c - static_cast<int>(q) * kUnit) - 1) << 14) |
                            (static_cast<uint32_t>(k) << 21);
        run_u += __shfl(su, 63, 64);
    }
    if (lane == 0) *s_nunits = run_u;
}

// One unit's records, one per lane (lane < cnt).
struct UnitRecords {
    int nanb;
    float4 q;     // q1, q2, s, Re v
    float vy;
};

// Position of the unit's footprint: tile-relative start (a0l, b0l) and, for preprocessed input,
// the tile origin in centred grid coordinates (the ES argument uses absolute positions there).
struct UnitPos {
    int a0l, b0l, oac, obc;
};

template <bool kSnap>
__device__ __forceinline__ float es_arg(int cell, float q, int a0l, int oac) {
    if constexpr (kSnap) return (static_cast<float>(cell) - q) * 0.25f;
    else return (static_cast<float>(oac + a0l + cell) - q) * 0.25f;
}

// Ninth row / column of one record's footprint (exact integer position, na or nb = 9), added
// straight to the LDS tile: lanes 0..8 take row 8 (columns 0..8), lanes 9..16 column 8 (rows
// 0..7).  The record's fields are wave-uniform.
template <bool kSnap, int T, int KMAX>
__device__ __forceinline__ void add_ninth(int na, int nb, float q1, float q2, float s, float vx, float vy,
                                          const UnitPos& p, int K, float* acc_re, float* acc_im) {
    const int lane = threadIdx.x & 63;
    int ei = -1, ej = -1;
    if (lane < 9 && na == 9) {
        ei = 8;
        ej = lane;
    } else if (lane >= 9 && lane < 17 && nb == 9) {
        ei = lane - 9;
        ej = 8;
    }
    if (ei < 0 || ej >= nb || ei >= na) return;
    const int ca = p.a0l + ei, cb = p.b0l + ej;
    if (ca < 0 || ca >= T || cb < 0 || cb >= T) return;
    const float ekv = es_kernel_fast(es_arg<kSnap>(ei, q1, p.a0l, p.oac)) *
                      es_kernel_fast(es_arg<kSnap>(ej, q2, p.b0l, p.obc));
    const float er = vx * ekv, eim = vy * ekv;
    float t2 = 1.0f, t1 = s;
    for (int k = 0; k < K; ++k) {
        float tk;
        if (k == 0) tk = 1.0f;
        else if (k == 1) tk = s;
        else {
            tk = 2.0f * s * t1 - t2;
            t2 = t1;
            t1 = tk;
        }
        const int idx = k * T * T + ca * T + swz<T>(ca, cb);
        atomicAdd(&acc_re[idx], tk * er);
        atomicAdd(&acc_im[idx], tk * eim);
    }
}

// FP32 unit accumulation: lane = footprint cell (8 x 8), the unit's records broadcast with
// readlane, 2K FP32 accumulators per lane, one LDS add per cell and term at the end.
template <bool kSnap, int T, int KMAX>
__device__ __forceinline__ void unit_fp32(const UnitRecords& m, int cnt, const UnitPos& p, int K,
                                          float* acc_re, float* acc_im) {
    const int lane = threadIdx.x & 63;
    const int ci = lane >> 3, cj = lane & 7;
    float re[KMAX], im[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) re[k] = im[k] = 0.0f;
    for (int r = 0; r < cnt; ++r) {
        const int nanb = readlane_i(m.nanb, r);
        const float q1 = readlane_f(m.q.x, r), q2 = readlane_f(m.q.y, r);
        const float s = readlane_f(m.q.z, r);
        const float vx = readlane_f(m.q.w, r), vy = readlane_f(m.vy, r);
        const int na = nanb & 15, nb = nanb >> 4;
        const float ka = es_kernel_fast(es_arg<kSnap>(ci, q1, p.a0l, p.oac));
        const float kb = es_kernel_fast(es_arg<kSnap>(cj, q2, p.b0l, p.obc));
        const float kv = (ci < na ? ka : 0.0f) * (cj < nb ? kb : 0.0f);
        const float wr = vx * kv, wi = vy * kv;
        float tkm2 = 1.0f, tkm1 = s;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (k < K) {
                float tk;
                if (k == 0) tk = 1.0f;
                else if (k == 1) tk = s;
                else {
                    tk = 2.0f * s * tkm1 - tkm2;
                    tkm2 = tkm1;
                    tkm1 = tk;
                }
                re[k] += tk * wr;
                im[k] += tk * wi;
            }
        }
        if (na == 9 || nb == 9) add_ninth<kSnap, T, KMAX>(na, nb, q1, q2, s, vx, vy, p, K, acc_re, acc_im);
    }
    const int ca = p.a0l + ci, cb = p.b0l + cj;
    if (ca >= 0 && ca < T && cb >= 0 && cb < T) {
        const int cell = ca * T + swz<T>(ca, cb);
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (k < K) {
                atomicAdd(&acc_re[k * T * T + cell], re[k]);
                atomicAdd(&acc_im[k * T * T + cell], im[k]);
            }
        }
    }
}

__device__ __forceinline__ void split_hilo(const float4& x0, const float4& x1, half8& hi, half8& lo) {
    const float v[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const _Float16 h = static_cast<_Float16>(v[i]);
        hi[i] = h;
        lo[i] = static_cast<_Float16>(v[i] - static_cast<float>(h));
    }
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MFMA unit accumulation.  All records of a unit share the footprint position, so the unit's
// contribution to its 8 x 8 cells is one contraction over records:
//   D[m][cell] = sum_r c[m][r] * B[r][cell],  m = 2k + ri,
//   c[2k + ri][r] = T_k(s_r) * (Re | Im) v_r,  B[r][(i, j)] = ES(x_r,i) ES(y_r,j),
// run as v_mfma_f32_16x16x32_f16 over chunks of 32 records (M = 16 rows of m per block, N = 16
// cells per block, 4 blocks for the 64 cells) with FP32 accumulation.  Each FP32 operand is
// split into f16 hi + lo and the product taken as hi*hi + lo*hi + hi*lo, which keeps ~22
// significant bits.  c is scaled by a power of two per unit (its largest |v| to ~2^14) and the
// ES factors by 2^7 each, so nothing overflows f16 and little falls into its subnormals; the
// scales are undone exactly when the unit is added to the LDS tile.
template <bool kSnap, int T, int KMAX>
__device__ __forceinline__ void unit_mfma(const UnitRecords& m, int cnt, const UnitPos& p, int K,
                                          float* acc_re, float* acc_im, float* st) {
    constexpr int MB = KMAX / 8;                   // 16-row blocks of m = 2k + ri
    const int lane = threadIdx.x & 63;
    float* s_ku = st;
    float* s_kv = st + 8 * kStageStride;
    float* s_c = st + 16 * kStageStride;

    // Unit scale from the largest |v| (a power of two, exact to apply and undo).
    float vmax = lane < cnt ? fmaxf(fabsf(m.q.w), fabsf(m.vy)) : 0.0f;
    for (int d = 32; d >= 1; d >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, d, 64));
    int e = 0;
    if (vmax > 0.0f && vmax <= 3.0e38f) (void)frexpf(vmax, &e);
    e = min(max(e, -100), 100);
    const float vscale = ldexpf(1.0f, 14 - e);

    // Ninth rows / columns (records at exact integer positions) go straight to LDS.
    uint64_t ninth = __ballot(lane < cnt && ((m.nanb & 15) == 9 || (m.nanb >> 4) == 9));
    while (ninth) {
        const int r = __builtin_ctzll(ninth);
        ninth &= ninth - 1;
        const int nanb = readlane_i(m.nanb, r);
        add_ninth<kSnap, T, KMAX>(nanb & 15, nanb >> 4, readlane_f(m.q.x, r), readlane_f(m.q.y, r),
                                  readlane_f(m.q.z, r), readlane_f(m.q.w, r), readlane_f(m.vy, r), p, K,
                                  acc_re, acc_im);
    }

    floatx4 d[MB][4];
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) d[b][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int rl = lane & 31, axis = lane >> 5;    // staging: lane = (record, axis)
    const int g8 = 8 * (lane >> 4), n16 = lane & 15;
    for (int c0 = 0; c0 < cnt; c0 += kChunk) {
        // Stage 32 records: the record's fields come from lane c0 + rl.
        const int src = c0 + rl;
        const int nanb = __shfl(m.nanb, src, 64);
        const float qa = __shfl(axis ? m.q.y : m.q.x, src, 64);
        const float s = __shfl(m.q.z, src, 64);
        const float v = __shfl(axis ? m.vy : m.q.w, src, 64);
        const bool valid = src < cnt;
        const int n_ax = axis ? (nanb >> 4) : (nanb & 15);
        const int l0 = axis ? p.b0l : p.a0l, o0 = axis ? p.obc : p.oac;
        float* s_k = axis ? s_kv : s_ku;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            float w = 0.0f;
            if (valid && i < n_ax) {
                float x;
                if constexpr (kSnap) x = (static_cast<float>(i) - qa) * 0.25f;
                else x = (static_cast<float>(o0 + l0 + i) - qa) * 0.25f;
                w = es_kernel_fast(x) * kEsScale;
            }
            s_k[i * kStageStride + rl] = w;
        }
        const float vs = valid ? v * vscale : 0.0f;
        float tkm2 = 1.0f, tkm1 = s;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            float tk;
            if (k == 0) tk = 1.0f;
            else if (k == 1) tk = s;
            else {
                tk = 2.0f * s * tkm1 - tkm2;
                tkm2 = tkm1;
                tkm1 = tk;
            }
            s_c[(2 * k + axis) * kStageStride + rl] = k < K ? tk * vs : 0.0f;
        }
        wave_lds_sync();

        // Operands.  A (16 x 32): lane holds A[n16][g8 + i]; B (32 x 16): lane holds
        // B[g8 + i][n16]; D (16 x 16): lane holds D[4 (lane / 16) + i][n16].
        half8 a_hi[MB], a_lo[MB];
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const float* row = s_c + (16 * b + n16) * kStageStride + g8;
            split_hilo(*reinterpret_cast<const float4*>(row), *reinterpret_cast<const float4*>(row + 4),
                       a_hi[b], a_lo[b]);
        }
        const float* kvrow = s_kv + (n16 & 7) * kStageStride + g8;
        const float4 kv0 = *reinterpret_cast<const float4*>(kvrow);
        const float4 kv1 = *reinterpret_cast<const float4*>(kvrow + 4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float* kurow = s_ku + (2 * j + (n16 >> 3)) * kStageStride + g8;
            const float4 ku0 = *reinterpret_cast<const float4*>(kurow);
            const float4 ku1 = *reinterpret_cast<const float4*>(kurow + 4);
            const float4 x0 = make_float4(ku0.x * kv0.x, ku0.y * kv0.y, ku0.z * kv0.z, ku0.w * kv0.w);
            const float4 x1 = make_float4(ku1.x * kv1.x, ku1.y * kv1.y, ku1.z * kv1.z, ku1.w * kv1.w);
            half8 b_hi, b_lo;
            split_hilo(x0, x1, b_hi, b_lo);
#pragma unroll
            for (int b = 0; b < MB; ++b) {
                d[b][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi[b], b_hi, d[b][j], 0, 0, 0);
                d[b][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_lo[b], b_hi, d[b][j], 0, 0, 0);
                d[b][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi[b], b_lo, d[b][j], 0, 0, 0);
            }
        }
        wave_lds_sync();
    }

    // Add the unit to the LDS tile, undoing the scales.
    const float inv = ldexpf(1.0f, e - 14 - kEsScaleLog2Sq);
#pragma unroll
    for (int b = 0; b < MB; ++b) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ca = p.a0l + 2 * j + (n16 >> 3), cb = p.b0l + (n16 & 7);
            if (ca < 0 || ca >= T || cb < 0 || cb >= T) continue;
            const int cell = ca * T + swz<T>(ca, cb);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int mm = 16 * b + 4 * (lane >> 4) + i;
                const int k = mm >> 1;
                if (k < K) atomicAdd(&((mm & 1) ? acc_im : acc_re)[k * T * T + cell], d[b][j][i] * inv);
            }
        }
    }
}

// Persistent workgroups take items from a queue.  An item's records are sorted by footprint
// position; its units (runs of at most 64 same-position records) are listed from the bucket
// offsets and shared out over the wavefronts.  Units of at least `gp.mfma_min` records take the
// MFMA contraction (kMfma), the others the FP32 broadcast loop.  The tile is then stored to the
// planes with its checkerboard sign, or added atomically when the tile is split over items.
template <bool kSnap, int T, int KMAX, int WAVES, bool kMfma>
__global__ void __launch_bounds__(WAVES * 64)
k_tile_grid(const Item* __restrict__ items, uint32_t* __restrict__ totals,
            const uint32_t* __restrict__ boff, Records rec, GridParams gp, TileGeom tg,
            float2* __restrict__ planes) {
    using L = TileLds<T, KMAX, WAVES, kMfma>;
    constexpr int KS = L::KS, NKEYS = L::NKEYS, NT = WAVES * 64;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* acc_re = reinterpret_cast<float*>(lds_raw + L::acc);
    float* acc_im = acc_re + KMAX * T * T;
    uint32_t* units = reinterpret_cast<uint32_t*>(lds_raw + L::units);
    __shared__ uint32_t s_nunits;
    __shared__ int s_item;

    const int K = gp.K;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* stage = reinterpret_cast<float*>(lds_raw + L::stage) + wave * L::stage_floats;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = static_cast<int>(atomicAdd(&totals[2], 1u));
        __syncthreads();
        if (s_item >= static_cast<int>(totals[1])) break;
        const Item it = items[s_item];
        const int ls = it.bin / tg.ntiles, tile = it.bin - ls * tg.ntiles;
        const int ta = tile / tg.ntx, tb = tile - ta * tg.ntx;
        const int off = gp.G / 2;
        for (int e = threadIdx.x; e < 2 * KMAX * T * T; e += NT) acc_re[e] = 0.0f;
        if (threadIdx.x < 64) build_units<NKEYS>(it, boff, units, &s_nunits);
        __syncthreads();

        const float4* q_g = rec.q + it.start;
        const float* vy_g = rec.vy + it.start;
        const uint8_t* nanb_g = rec.nanb + it.start;
        const uint32_t nunits = s_nunits;
        for (uint32_t u = wave; u < nunits; u += WAVES) {
            const uint32_t w = units[u];
            const int start = w & 16383, cnt = ((w >> 14) & 127) + 1, key = w >> 21;
            UnitPos p;
            p.a0l = key / KS - 8;
            p.b0l = key % KS - 8;
            p.oac = ta * T - off;                         // tile origin, centred coordinates
            p.obc = tb * T - off;
            UnitRecords m;
            m.nanb = 0;
            m.q = make_float4(0.f, 0.f, 0.f, 0.f);
            m.vy = 0.f;
            if (lane < cnt) {                             // lane l loads record l of the unit
                m.nanb = nanb_g[start + lane];
                m.q = q_g[start + lane];
                m.vy = vy_g[start + lane];
            }
            if constexpr (kMfma) {
                if (cnt >= gp.mfma_min) {
                    unit_mfma<kSnap, T, KMAX>(m, cnt, p, K, acc_re, acc_im, stage);
                    continue;
                }
            }
            unit_fp32<kSnap, T, KMAX>(m, cnt, p, K, acc_re, acc_im);
        }
        __syncthreads();

        const int G = gp.G;
        const size_t plane_size = static_cast<size_t>(G) * G;
        for (int e = threadIdx.x; e < K * T * T; e += NT) {
            const int k = e / (T * T), c = e - k * T * T;
            const int a = c / T, b = c - a * T;
            const int ag = ta * T + a, bg = tb * T + b;
            if (ag >= G || bg >= G) continue;
            const int lc = k * T * T + a * T + swz<T>(a, b);
            float2 val = make_float2(acc_re[lc], acc_im[lc]);
            if ((ag + bg) & 1) {
                val.x = -val.x;
                val.y = -val.y;
            }
            float2* dst = planes + (static_cast<size_t>(ls) * K + k) * plane_size + static_cast<size_t>(ag) * G + bg;
            if (it.split) {
                unsafeAtomicAdd(&dst->x, val.x);
                unsafeAtomicAdd(&dst->y, val.y);
            } else {
                *dst = val;
            }
        }
    }
}

template <bool kSnap, int T, int KMAX, int WAVES, bool kMfma>
void launch_tile_kernel(svdi_context* ctx, const Item* items, uint32_t* totals, const uint32_t* boff,
                        const Records& rec, const GridParams& gp, const TileGeom& tg, float2* planes) {
    using L = TileLds<T, KMAX, WAVES, kMfma>;
    constexpr size_t bytes = L::bytes;
    static_assert(bytes <= 160 * 1024 - 64, "tile LDS exceeds a CU");
    if (tg.T != T || tg.KS != L::KS || gp.K > KMAX)
        throw std::logic_error("tile kernel instantiated for another tile geometry");
    SVDI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_grid<kSnap, T, KMAX, WAVES, kMfma>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / (bytes + 64)));
    const unsigned grid = static_cast<unsigned>(std::min(per_cu, 2) * std::max(1, ctx->cu_count));
    k_tile_grid<kSnap, T, KMAX, WAVES, kMfma><<<grid, WAVES * 64, bytes, ctx->stream>>>(items, totals, boff, rec,
                                                                                      gp, tg, planes);
    SVDI_LAUNCH_CHECK();
}

template <bool kSnap, bool kMfma>
void launch_tile_dispatch(svdi_context* ctx, const Item* items, uint32_t* totals, const uint32_t* boff,
                          const Records& rec, const GridParams& gp, const TileGeom& tg, float2* planes) {
    if constexpr (kMfma) {
        if (gp.K <= 8) launch_tile_kernel<kSnap, 32, 8, 16, true>(ctx, items, totals, boff, rec, gp, tg, planes);
        else if (gp.K <= 16) launch_tile_kernel<kSnap, 16, 16, 8, true>(ctx, items, totals, boff, rec, gp, tg, planes);
        else launch_tile_kernel<kSnap, 16, 32, 4, true>(ctx, items, totals, boff, rec, gp, tg, planes);
    } else {
        if (gp.K <= 4) launch_tile_kernel<kSnap, 32, 4, 8, false>(ctx, items, totals, boff, rec, gp, tg, planes);
        else if (gp.K <= 8) launch_tile_kernel<kSnap, 32, 8, 8, false>(ctx, items, totals, boff, rec, gp, tg, planes);
        else if (gp.K <= 16) launch_tile_kernel<kSnap, 16, 16, 8, false>(ctx, items, totals, boff, rec, gp, tg, planes);
        else launch_tile_kernel<kSnap, 16, 32, 8, false>(ctx, items, totals, boff, rec, gp, tg, planes);
    }
}

}  // namespace

const uint32_t* launch_grid_tiled(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                                  float2* planes, bool mfma) {
    if (gp.K < 1 || gp.K > kMaxTerms) throw std::logic_error("term count out of range");
    TileGeom tg;
    tg.T = gp.K <= 8 ? 32 : 16;
    tg.ntx = (gp.G + tg.T - 1) / tg.T;
    tg.ntiles = tg.ntx * tg.ntx;
    tg.nbins = tg.ntiles * gp.nslabs;
    tg.KS = tg.T + 8;
    tg.nkeys = tg.KS * tg.KS;
    const size_t nbuckets = static_cast<size_t>(tg.nbins) * tg.nkeys;
    const size_t S = in.num_samples;
    const size_t max_records = 4 * S;                      // a footprint overlaps at most 2 x 2 tiles
    if (max_records >= (1ull << 31) || nbuckets >= (1ull << 31))
        throw UnsupportedError("problem too large for the tiled gridder");
    tg.nbuckets = static_cast<uint32_t>(nbuckets);
    const size_t max_items = static_cast<size_t>(tg.nbins) + (max_records + kMinItem - 1) / kMinItem;
    const size_t nscan = (nbuckets + kScanBlock - 1) / kScanBlock;

    // bcount[0, nbuckets): counts; bcount[nbuckets + 2, +6): totals.  boff[nbuckets] = total.
    uint32_t* bcount = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinCounts, (nbuckets + 8) * sizeof(uint32_t)));
    uint32_t* boff = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinOffsets, (nbuckets + 8) * sizeof(uint32_t)));
    uint32_t* bcursor = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinCursor, nbuckets * sizeof(uint32_t)));
    uint32_t* scan_sums = static_cast<uint32_t*>(ctx->ws.get(Workspace::kScanSums, nscan * sizeof(uint32_t)));
    uint32_t* totals = bcount + nbuckets + 2;
    Item* items = static_cast<Item*>(ctx->ws.get(Workspace::kItems, max_items * sizeof(Item)));
    Records rec;
    rec.q = static_cast<float4*>(ctx->ws.get(Workspace::kRecPos, max_records * sizeof(float4)));
    char* aux = static_cast<char*>(ctx->ws.get(Workspace::kRecVim, max_records * (sizeof(float) + 1)));
    rec.vy = reinterpret_cast<float*>(aux);
    rec.nanb = reinterpret_cast<uint8_t*>(aux + max_records * sizeof(float));

    launch_fill(ctx, reinterpret_cast<float*>(bcount), nbuckets + 8, 0.0f);
    unsigned bin_blocks;
    if (in.from_snapshot) bin_blocks = static_cast<unsigned>((in.num_rows + kBinThreads - 1) / kBinThreads);
    else bin_blocks = static_cast<unsigned>((S + kBinThreads * kBinPerThread - 1) / (kBinThreads * kBinPerThread));
    if (in.from_snapshot) k_bin_count<true><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, bcount);
    else k_bin_count<false><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, bcount);
    SVDI_LAUNCH_CHECK();
    const uint32_t n32 = static_cast<uint32_t>(nbuckets);
    k_scan_blocks<<<static_cast<unsigned>(nscan), 1024, 0, ctx->stream>>>(bcount, boff, scan_sums, n32);
    SVDI_LAUNCH_CHECK();
    k_scan_sums<<<1, 1024, 0, ctx->stream>>>(scan_sums, static_cast<int>(nscan), boff + nbuckets);
    SVDI_LAUNCH_CHECK();
    k_scan_add<<<static_cast<unsigned>((nbuckets + 255) / 256), 256, 0, ctx->stream>>>(boff, scan_sums, n32, bcursor);
    SVDI_LAUNCH_CHECK();
    const int cus = std::max(1, ctx->cu_count);
    k_make_items<<<1, 1024, 0, ctx->stream>>>(boff, tg, items, static_cast<int>(max_items), cus, totals);
    SVDI_LAUNCH_CHECK();
    k_zero_split<<<tg.nbins, 256, 0, ctx->stream>>>(boff, totals, tg, gp.G, gp.K, planes);
    SVDI_LAUNCH_CHECK();
    if (in.from_snapshot)
        k_bin_scatter<true><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, bcursor, rec);
    else
        k_bin_scatter<false><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, bcursor, rec);
    SVDI_LAUNCH_CHECK();
    const size_t tok = stage_begin(ctx, 5);
    if (mfma) {
        if (in.from_snapshot) launch_tile_dispatch<true, true>(ctx, items, totals, boff, rec, gp, tg, planes);
        else launch_tile_dispatch<false, true>(ctx, items, totals, boff, rec, gp, tg, planes);
    } else {
        if (in.from_snapshot) launch_tile_dispatch<true, false>(ctx, items, totals, boff, rec, gp, tg, planes);
        else launch_tile_dispatch<false, false>(ctx, items, totals, boff, rec, gp, tg, planes);
    }
    stage_end(ctx, tok);
    return totals;
}

}  // namespace svdi
