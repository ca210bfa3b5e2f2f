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
#include "scan.hpp"

namespace svdi {
namespace {

constexpr int kBinThreads = 256;
constexpr int kBinPerThread = 64;        // samples per thread, preprocessed input
constexpr int kMinItem = 4096;           // smallest cap on the records of one work item
constexpr int kMaxItem = 16384;          // largest cap
constexpr int kTileThreads = 512;
constexpr int kUnit = 64;                // records per accumulation unit (one wavefront)

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

// Snapshot input, one thread per row.  A short baseline keeps its footprint position over many
// consecutive channels, and the densest buckets (~1e5 records at the uv centre) are made of such
// runs, so the count pass takes one atomic per run and overlapped tile, not per sample.  A run
// spans at most 64 channels and holds at most 64 samples.
//   count:   each accepted sample stores its place in its run, runpos = d | i << 6 (d = channel
//            offset from the run's first channel, i = index among the run's samples); at the
//            run's end, per overlapped tile t, the run's offset inside the bucket (the value the
//            atomic add returned) goes to runoff[t][first sample of the run];
//   scatter: slot = boff[bucket] + runoff[t][first sample] + i, with no atomics, and every
//            sample handled in its own loop iteration (no divergent per-run loops).
constexpr int kMaxRun = 64;

template <class Fn>
__device__ __forceinline__ void for_tiles(const Footprint& f, const TileGeom& tg, Fn&& fn) {
    int t = 0;
    for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
        for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb) fn(t++, ta, tb);
}

template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_bin_count(SampleInput in, GridParams gp, TileGeom tg, uint32_t* __restrict__ bcount,
            uint16_t* __restrict__ runpos, uint32_t* __restrict__ runoff) {
    take_device_plan(gp);
    if constexpr (kSnap) {
        const size_t row = blockIdx.x * (size_t)kBinThreads + threadIdx.x;
        if (row >= in.num_rows) return;
        const size_t R = in.num_rows, S = in.num_samples;
        Footprint run;
        uint32_t len = 0;
        size_t begin = 0;
        auto close_run = [&]() {
            const size_t first = row + begin * R;
            for_tiles(run, tg, [&](int t, int ta, int tb) {
                runoff[t * S + first] = atomicAdd(&bcount[bucket_of(run, tg, ta, tb)], len);
            });
        };
        for (size_t chan = 0; chan < in.num_channels; ++chan) {
            const size_t idx = row + chan * R;
            Footprint f;
            if (!setup_sample<true>(in, gp, idx, row, chan, f)) continue;
            if (len > 0 && (!same_footprint(f, run) || chan - begin >= kMaxRun || len >= kMaxRun)) {
                close_run();
                len = 0;
            }
            if (len == 0) {
                run = f;
                begin = chan;
            }
            runpos[idx] = static_cast<uint16_t>((chan - begin) | (len << 6));
            ++len;
        }
        if (len > 0) close_run();
    } else {
        for_sample_tiles<false>(in, gp, tg, [&](const Footprint& f, int ta, int tb) {
            atomicAdd(&bcount[bucket_of(f, tg, ta, tb)], 1u);
        });
    }
}

// Records are tile-independent (positions relative to the footprint start or absolute), so a
// sample overlapping several tiles writes the same record into each tile's bucket.  `limit`
// (the record buffer's capacity) only bounds the writes.
template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_bin_scatter(SampleInput in, GridParams gp, TileGeom tg, const uint32_t* __restrict__ boff,
              uint32_t* __restrict__ bcursor, const uint16_t* __restrict__ runpos,
              const uint32_t* __restrict__ runoff, Records rec, uint32_t limit) {
    take_device_plan(gp);
    if constexpr (kSnap) {
        const size_t row = blockIdx.x * (size_t)kBinThreads + threadIdx.x;
        if (row >= in.num_rows) return;
        const size_t R = in.num_rows, S = in.num_samples;
        for (size_t chan = 0; chan < in.num_channels; ++chan) {
            const size_t idx = row + chan * R;
            Footprint f;
            if (!setup_sample<true>(in, gp, idx, row, chan, f)) continue;
            const uint32_t rp = runpos[idx];
            const size_t first = row + (chan - (rp & 63)) * R;
            const uint32_t i = rp >> 6;
            for_tiles(f, tg, [&](int t, int ta, int tb) {
                const uint32_t slot = boff[bucket_of(f, tg, ta, tb)] + runoff[t * S + first] + i;
                if (slot < limit) write_record(rec, slot, f);
            });
        }
    } else {
        for_sample_tiles<false>(in, gp, tg, [&](const Footprint& f, int ta, int tb) {
            const uint32_t slot = atomicAdd(&bcursor[bucket_of(f, tg, ta, tb)], 1u);
            if (slot < limit) write_record(rec, slot, f);
        });
    }
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

// LDS of the tile kernel: the K planes (re, im) of one T x T tile and the unit list of the
// current item.
template <int T, int KMAX>
struct TileLds {
    static constexpr int KS = T + 8;               // footprint starts in [-8, T) per axis
    static constexpr int NKEYS = KS * KS;
    static constexpr int MAXUNITS = NKEYS + kMaxItem / kUnit;
    static constexpr size_t acc = 0;                                        // float [2][KMAX][T*T]
    static constexpr size_t units = acc + sizeof(float) * 2 * KMAX * T * T;  // u32 [MAXUNITS]
    static constexpr size_t bytes = units + sizeof(uint32_t) * MAXUNITS;
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

// Persistent workgroups take items from a queue.  An item's records are sorted by footprint
// position; its units (runs of at most 64 same-position records) are listed from the bucket
// offsets and shared out over the wavefronts, each accumulated by the FP32 broadcast loop.  The
// tile is then stored to the
// planes with its checkerboard sign, or added atomically when the tile is split over items.
template <bool kSnap, int T, int KMAX, int WAVES>
__global__ void __launch_bounds__(WAVES * 64)
k_tile_grid(const Item* __restrict__ items, uint32_t* __restrict__ totals,
            const uint32_t* __restrict__ boff, Records rec, GridParams gp, TileGeom tg,
            float2* __restrict__ planes) {
    using L = TileLds<T, KMAX>;
    constexpr int KS = L::KS, NKEYS = L::NKEYS, NT = WAVES * 64;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* acc_re = reinterpret_cast<float*>(lds_raw + L::acc);
    float* acc_im = acc_re + KMAX * T * T;
    uint32_t* units = reinterpret_cast<uint32_t*>(lds_raw + L::units);
    __shared__ uint32_t s_nunits;
    __shared__ int s_item;

    const int K = gp.K;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
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

template <bool kSnap, int T, int KMAX, int WAVES>
void launch_tile_kernel(svdi_context* ctx, const Item* items, uint32_t* totals, const uint32_t* boff,
                        const Records& rec, const GridParams& gp, const TileGeom& tg, float2* planes) {
    using L = TileLds<T, KMAX>;
    constexpr size_t bytes = L::bytes;
    static_assert(bytes <= 160 * 1024 - 64, "tile LDS exceeds a CU");
    if (tg.T != T || tg.KS != L::KS || gp.K > KMAX)
        throw std::logic_error("tile kernel instantiated for another tile geometry");
    allow_dynamic_lds(reinterpret_cast<const void*>(&k_tile_grid<kSnap, T, KMAX, WAVES>), bytes);
    const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / (bytes + 64)));
    const unsigned grid = static_cast<unsigned>(std::min(per_cu, 2) * std::max(1, ctx->cu_count));
    k_tile_grid<kSnap, T, KMAX, WAVES><<<grid, WAVES * 64, bytes, ctx->stream>>>(items, totals, boff, rec,
                                                                                      gp, tg, planes);
    SVDI_LAUNCH_CHECK();
}

template <bool kSnap>
void launch_tile_dispatch(svdi_context* ctx, const Item* items, uint32_t* totals, const uint32_t* boff,
                          const Records& rec, const GridParams& gp, const TileGeom& tg, float2* planes) {
    if (gp.K <= 4) launch_tile_kernel<kSnap, 32, 4, 8>(ctx, items, totals, boff, rec, gp, tg, planes);
    else if (gp.K <= 8) launch_tile_kernel<kSnap, 32, 8, 8>(ctx, items, totals, boff, rec, gp, tg, planes);
    else if (gp.K <= 16) launch_tile_kernel<kSnap, 16, 16, 8>(ctx, items, totals, boff, rec, gp, tg, planes);
    else launch_tile_kernel<kSnap, 16, 32, 8>(ctx, items, totals, boff, rec, gp, tg, planes);
}

}  // namespace

const uint32_t* launch_grid_tiled(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                                  float2* planes) {
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
    const size_t nscan = scan_blocks(nbuckets);

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
    uint16_t* runpos = nullptr;
    uint32_t* runoff = nullptr;
    if (in.from_snapshot) {
        runpos = static_cast<uint16_t*>(ctx->ws.get(Workspace::kRunPos, S * sizeof(uint16_t)));
        runoff = static_cast<uint32_t*>(ctx->ws.get(Workspace::kRunOff, 4 * S * sizeof(uint32_t)));
        k_bin_count<true><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, bcount, runpos, runoff);
    } else {
        k_bin_count<false><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, bcount, runpos, runoff);
    }
    SVDI_LAUNCH_CHECK();
    launch_exclusive_scan(ctx, bcount, boff, bcursor, scan_sums, nbuckets);
    const int cus = std::max(1, ctx->cu_count);
    k_make_items<<<1, 1024, 0, ctx->stream>>>(boff, tg, items, static_cast<int>(max_items), cus, totals);
    SVDI_LAUNCH_CHECK();
    k_zero_split<<<tg.nbins, 256, 0, ctx->stream>>>(boff, totals, tg, gp.G, gp.K, planes);
    SVDI_LAUNCH_CHECK();
    const uint32_t limit = static_cast<uint32_t>(max_records);
    if (in.from_snapshot)
        k_bin_scatter<true><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, boff, bcursor, runpos,
                                                                         runoff, rec, limit);
    else
        k_bin_scatter<false><<<bin_blocks, kBinThreads, 0, ctx->stream>>>(in, gp, tg, boff, bcursor, runpos,
                                                                          runoff, rec, limit);
    SVDI_LAUNCH_CHECK();
    const size_t tok = stage_begin(ctx, 5);
    if (in.from_snapshot) launch_tile_dispatch<true>(ctx, items, totals, boff, rec, gp, tg, planes);
    else launch_tile_dispatch<false>(ctx, items, totals, boff, rec, gp, tg, planes);
    stage_end(ctx, tok);
    return totals;
}

}  // namespace svdi
