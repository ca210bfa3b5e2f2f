// Tiled ES gridder: samples binned into T x T grid tiles, each tile accumulated in LDS.
//
// Same sample model as the reference's fused_gridding (SVDIntegratedImager.cu:1218-1293), with
// a different schedule.  The reference adds every (cell, term) contribution to the G x G planes
// with a device atomic; here:
//   1. count:   each sample's clipped footprint is assigned to every tile it overlaps (a sample
//               straddling a tile edge is duplicated, so tiles own their cells exclusively);
//   2. scan:    per-bin record offsets and work items of similar size (hot tiles in the dense
//               uv core are split into several items);
//   3. scatter: per-(sample, tile) records written contiguously per bin;
//   4. grid:    a persistent workgroup per CU takes items from a queue, counting-sorts each
//               item's records by exact footprint position, and lets each wavefront accumulate
//               a run of same-position records in registers (lane = footprint cell) before one
//               LDS add per cell and term; the tile is then stored to the planes with its
//               checkerboard sign (split tiles add atomically into a zeroed tile instead).
// Bins are (local slab, tile); the planes need no clearing except split tiles.
#include "sample.hpp"

namespace svdi {
namespace {

constexpr int kBinThreads = 256;
constexpr int kBinPerThread = 64;        // samples per thread, preprocessed input
constexpr int kLdsBins = 16384;          // LDS histogram capacity of the binning passes
constexpr int kSub = 1024;               // records staged in LDS at a time
constexpr int kMinItem = 4 * kSub;       // smallest cap on the records of one work item
constexpr int kMaxItem = 16384;          // largest cap (the per-item order array is 16-bit)
constexpr int kTileThreads = 512;
constexpr int kUnit = 64;                // records per accumulation unit (one wavefront)

struct TileGeom {
    int T, ntx, ntiles, nbins, use_lds;
};

struct Item {
    int bin, start, count, split;
};

// Records, structure of arrays.  ab: footprint relative to the record's tile, bits [0,6)
// a0 + 8, [6,12) b0 + 8, [12,16) rows na, [16,20) columns nb; q = (q1, q2, s, Re v); vy = Im v.
struct Records {
    uint32_t* ab;
    float4* q;
    float* vy;
};

__device__ __forceinline__ int pack_ab(int a0l, int b0l, int na, int nb) {
    return (a0l + 8) | ((b0l + 8) << 6) | (na << 12) | (nb << 16);
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

template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_bin_count(SampleInput in, GridParams gp, TileGeom tg, uint32_t* __restrict__ counts) {
    extern __shared__ uint32_t hist[];
    if (tg.use_lds) {
        for (int i = threadIdx.x; i < tg.nbins; i += kBinThreads) hist[i] = 0;
        __syncthreads();
    }
    for_block_samples<kSnap>(in, [&](size_t idx, size_t row, size_t chan) {
        Footprint f;
        if (!setup_sample<kSnap>(in, gp, idx, row, chan, f)) return;
        const int base = f.local_slab * tg.ntiles;
        for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
            for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb) {
                const int bin = base + ta * tg.ntx + tb;
                if (tg.use_lds) atomicAdd(&hist[bin], 1u);
                else atomicAdd(&counts[bin], 1u);
            }
    });
    if (tg.use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < tg.nbins; i += kBinThreads)
            if (hist[i]) atomicAdd(&counts[i], hist[i]);
    }
}

// One block: exclusive scan of the bin counts into record offsets, then the work items.  An
// item holds at most `cap` records, cap = max(kMinItem, total / (4 x CUs)) rounded to kSub, so
// the dense uv-core tiles split into a few items of similar size (each item adds its tile
// back once) and every bin has at least one item (every tile is written).
// totals[0] = records, [1] = items, [2] = queue head of the tile kernel, [3] = cap.
__global__ void __launch_bounds__(1024)
k_bin_scan(const uint32_t* __restrict__ counts, int nbins, uint32_t* __restrict__ cursor,
           Item* __restrict__ items, int max_items, int cu_count, uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_a[1024];
    __shared__ uint32_t s_cap;
    const int tid = threadIdx.x;
    const int per = (nbins + 1023) / 1024;
    const int lo = min(tid * per, nbins), hi = min(lo + per, nbins);
    auto block_scan = [&](uint32_t v) {            // inclusive scan of v over the block
        s_a[tid] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t t = tid >= off ? s_a[tid - off] : 0u;
            __syncthreads();
            s_a[tid] += t;
            __syncthreads();
        }
        const uint32_t r = s_a[tid];
        __syncthreads();
        return r;
    };
    uint32_t rec = 0;
    for (int b = lo; b < hi; ++b) rec += counts[b];
    const uint32_t rec_incl = block_scan(rec);
    if (tid == 1023) {
        const uint32_t total = rec_incl;
        const uint32_t fair = (total + 4u * cu_count - 1) / (4u * cu_count);
        const uint32_t cap = min(static_cast<uint32_t>(kMaxItem), max(static_cast<uint32_t>(kMinItem), fair));
        s_cap = (cap + kSub - 1) / kSub * kSub;
        totals[0] = total;
        totals[3] = s_cap;
    }
    __syncthreads();
    const uint32_t cap = s_cap;
    uint32_t it = 0;
    for (int b = lo; b < hi; ++b) it += max(1u, (counts[b] + cap - 1) / cap);
    const uint32_t it_incl = block_scan(it);
    uint32_t base_rec = rec_incl - rec, base_it = it_incl - it;
    for (int b = lo; b < hi; ++b) {
        const uint32_t c = counts[b];
        cursor[b] = base_rec;
        const uint32_t n = max(1u, (c + cap - 1) / cap);
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t slot = base_it + q;
            if (slot < static_cast<uint32_t>(max_items)) {
                Item t;
                t.bin = b;
                t.start = static_cast<int>(base_rec + q * cap);
                t.count = static_cast<int>(min(cap, c - q * cap));
                t.split = n > 1;
                items[slot] = t;
            }
        }
        base_rec += c;
        base_it += n;
    }
    if (tid == 1023) totals[1] = min(it_incl, static_cast<uint32_t>(max_items));
}

template <bool kSnap>
__global__ void __launch_bounds__(kBinThreads)
k_bin_scatter(SampleInput in, GridParams gp, TileGeom tg, uint32_t* __restrict__ cursor,
              Records rec) {
    extern __shared__ uint32_t hist[];
    auto emit = [&](const Footprint& f, int ta, int tb, uint32_t slot) {
        rec.ab[slot] = static_cast<uint32_t>(pack_ab(f.a0 - ta * tg.T, f.b0 - tb * tg.T, f.na, f.nb));
        rec.q[slot] = make_float4(f.q1, f.q2, f.s, f.vw.x);
        rec.vy[slot] = f.vw.y;
    };
    if (tg.use_lds) {
        for (int i = threadIdx.x; i < tg.nbins; i += kBinThreads) hist[i] = 0;
        __syncthreads();
        for_block_samples<kSnap>(in, [&](size_t idx, size_t row, size_t chan) {
            Footprint f;
            if (!setup_sample<kSnap>(in, gp, idx, row, chan, f)) return;
            const int base = f.local_slab * tg.ntiles;
            for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
                for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb)
                    atomicAdd(&hist[base + ta * tg.ntx + tb], 1u);
        });
        __syncthreads();
        for (int i = threadIdx.x; i < tg.nbins; i += kBinThreads)
            if (hist[i]) hist[i] = atomicAdd(&cursor[i], hist[i]);      // this block's base
        __syncthreads();
        for_block_samples<kSnap>(in, [&](size_t idx, size_t row, size_t chan) {
            Footprint f;
            if (!setup_sample<kSnap>(in, gp, idx, row, chan, f)) return;
            const int base = f.local_slab * tg.ntiles;
            for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
                for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb)
                    emit(f, ta, tb, atomicAdd(&hist[base + ta * tg.ntx + tb], 1u));
        });
    } else {
        for_block_samples<kSnap>(in, [&](size_t idx, size_t row, size_t chan) {
            Footprint f;
            if (!setup_sample<kSnap>(in, gp, idx, row, chan, f)) return;
            const int base = f.local_slab * tg.ntiles;
            for (int ta = f.a0 / tg.T; ta <= (f.a0 + f.na - 1) / tg.T; ++ta)
                for (int tb = f.b0 / tg.T; tb <= (f.b0 + f.nb - 1) / tg.T; ++tb)
                    emit(f, ta, tb, atomicAdd(&cursor[base + ta * tg.ntx + tb], 1u));
        });
    }
}

// Clear the tiles that are accumulated by several items (their items add atomically).
__global__ void __launch_bounds__(256)
k_zero_split(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ totals, TileGeom tg,
             int G, int K, float2* __restrict__ planes) {
    const int bin = blockIdx.x;
    if (counts[bin] <= totals[3]) return;
    const int ls = bin / tg.ntiles, tile = bin - ls * tg.ntiles;
    const int ta = tile / tg.ntx, tb = tile - ta * tg.ntx;
    const size_t plane_size = static_cast<size_t>(G) * G;
    for (int e = threadIdx.x; e < K * tg.T * tg.T; e += 256) {
        const int k = e / (tg.T * tg.T), c = e - k * tg.T * tg.T;
        const int a = ta * tg.T + c / tg.T, b = tb * tg.T + c % tg.T;
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

__device__ __forceinline__ int rec_key(uint32_t ab, int KS) {
    return static_cast<int>(ab & 63) * KS + static_cast<int>((ab >> 6) & 63);
}

// LDS layout of the tile kernel.
template <int T, int KMAX>
struct TileLds {
    static constexpr int KS = T + 8;               // footprint starts in [-8, T) per axis
    static constexpr int NKEYS = KS * KS;
    static constexpr int MAXUNITS = NKEYS + kMaxItem / kUnit;
    static constexpr size_t acc = 0;                                        // float [2][KMAX][T*T]
    static constexpr size_t khist = acc + sizeof(float) * 2 * KMAX * T * T;  // u32 [NKEYS]
    static constexpr size_t koff = khist + sizeof(uint32_t) * NKEYS;
    static constexpr size_t uoff = koff + sizeof(uint32_t) * NKEYS;
    static constexpr size_t units = uoff + sizeof(uint32_t) * NKEYS;        // u32 [MAXUNITS]
    static constexpr size_t order = units + sizeof(uint32_t) * MAXUNITS;    // u16 [kMaxItem]
    static constexpr size_t bytes = order + sizeof(uint16_t) * kMaxItem;
};

// One workgroup per CU, persistent: items come from a queue.  Per item, the T x T x K tile
// accumulator lives in LDS; the item's records are counting-sorted by exact footprint position
// (u16 order in LDS) and cut into units of at most 64 same-position records.  A wavefront
// accumulates one unit: lane = footprint cell, the unit's records gathered one per lane and
// broadcast with readlane, 2K FP32 accumulators per lane, one LDS add per cell and term at the
// end of the unit.
template <bool kSnap, int T, int KMAX>
__global__ void __launch_bounds__(kTileThreads)
k_tile_grid(const Item* __restrict__ items, uint32_t* __restrict__ totals, Records rec,
            GridParams gp, TileGeom tg, float2* __restrict__ planes) {
    using L = TileLds<T, KMAX>;
    constexpr int KS = L::KS, NKEYS = L::NKEYS;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* acc_re = reinterpret_cast<float*>(lds_raw + L::acc);
    float* acc_im = acc_re + KMAX * T * T;
    uint32_t* khist = reinterpret_cast<uint32_t*>(lds_raw + L::khist);
    uint32_t* koff = reinterpret_cast<uint32_t*>(lds_raw + L::koff);
    uint32_t* uoff = reinterpret_cast<uint32_t*>(lds_raw + L::uoff);
    uint32_t* units = reinterpret_cast<uint32_t*>(lds_raw + L::units);
    uint16_t* order = reinterpret_cast<uint16_t*>(lds_raw + L::order);
    __shared__ uint32_t s_nunits;
    __shared__ int s_item;

    const int K = gp.K;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ci = lane >> 3, cj = lane & 7;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = static_cast<int>(atomicAdd(&totals[2], 1u));
        __syncthreads();
        if (s_item >= static_cast<int>(totals[1])) break;
        const Item it = items[s_item];
        const int ls = it.bin / tg.ntiles, tile = it.bin - ls * tg.ntiles;
        const int ta = tile / tg.ntx, tb = tile - ta * tg.ntx;
        const int off = gp.G / 2;
        const int oac = ta * T - off, obc = tb * T - off;    // tile origin, centred coordinates
        const uint32_t* ab_g = rec.ab + it.start;
        const float4* q_g = rec.q + it.start;
        const float* vy_g = rec.vy + it.start;
        for (int e = threadIdx.x; e < 2 * KMAX * T * T; e += kTileThreads) acc_re[e] = 0.0f;
        for (int e = threadIdx.x; e < NKEYS; e += kTileThreads) khist[e] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < it.count; i += kTileThreads) atomicAdd(&khist[rec_key(ab_g[i], KS)], 1u);
        __syncthreads();
        // Exclusive scans of records and units per key, one wavefront.
        if (threadIdx.x < 64) {
            uint32_t run_r = 0, run_u = 0;
            for (int base = 0; base < NKEYS; base += 64) {
                const int k = base + lane;
                const uint32_t c = k < NKEYS ? khist[k] : 0u;
                const uint32_t u = (c + kUnit - 1) / kUnit;
                uint32_t sr = c, su = u;
                for (int d = 1; d < 64; d <<= 1) {
                    const uint32_t tr = __shfl_up(sr, d, 64), tu = __shfl_up(su, d, 64);
                    if (lane >= d) {
                        sr += tr;
                        su += tu;
                    }
                }
                if (k < NKEYS) {
                    koff[k] = run_r + sr - c;
                    uoff[k] = run_u + su - u;
                }
                run_r += __shfl(sr, 63, 64);
                run_u += __shfl(su, 63, 64);
            }
            if (lane == 0) s_nunits = run_u;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < NKEYS; k += kTileThreads) {
            const uint32_t c = khist[k];
            khist[k] = koff[k];
            for (uint32_t q = 0; q * kUnit < c; ++q)
                units[uoff[k] + q] = (koff[k] + q * kUnit) |
                                     ((min(static_cast<uint32_t>(kUnit), c - q * kUnit) - 1) << 14) |
                                     (static_cast<uint32_t>(k) << 21);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < it.count; i += kTileThreads)
            order[atomicAdd(&khist[rec_key(ab_g[i], KS)], 1u)] = static_cast<uint16_t>(i);
        __syncthreads();

        for (uint32_t u = wave; u < s_nunits; u += kTileThreads / 64) {
            const uint32_t w = units[u];
            const int start = w & 16383, cnt = ((w >> 14) & 127) + 1, key = w >> 21;
            const int a0l = key / KS - 8, b0l = key % KS - 8;
            int m_ab = 0;
            float4 m_q = make_float4(0.f, 0.f, 0.f, 0.f);
            float m_vy = 0.f;
            if (lane < cnt) {                             // lane l gathers record l of the unit
                const int ri = order[start + lane];
                m_ab = static_cast<int>(ab_g[ri]);
                m_q = q_g[ri];
                m_vy = vy_g[ri];
            }
            float re[KMAX], im[KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) re[k] = im[k] = 0.0f;
            for (int r = 0; r < cnt; ++r) {
                const int ab = readlane_i(m_ab, r);
                const float q1 = readlane_f(m_q.x, r), q2 = readlane_f(m_q.y, r);
                const float s = readlane_f(m_q.z, r);
                const float vx = readlane_f(m_q.w, r), vy = readlane_f(m_vy, r);
                const int na = (ab >> 12) & 15, nb = (ab >> 16) & 15;
                float xa, xb;
                if constexpr (kSnap) {
                    xa = (static_cast<float>(ci) - q1) * 0.25f;
                    xb = (static_cast<float>(cj) - q2) * 0.25f;
                } else {
                    xa = (static_cast<float>(oac + a0l + ci) - q1) * 0.25f;
                    xb = (static_cast<float>(obc + b0l + cj) - q2) * 0.25f;
                }
                const float ka = es_kernel_fast(xa), kb = es_kernel_fast(xb);
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
                if (na == 9 || nb == 9) {
                    // Ninth row / column of a footprint at an exact integer position:
                    // lanes 0..8 take row 8 (columns 0..8), lanes 9..16 column 8 (rows 0..7).
                    int ei = -1, ej = -1;
                    if (lane < 9 && na == 9) {
                        ei = 8;
                        ej = lane;
                    } else if (lane >= 9 && lane < 17 && nb == 9) {
                        ei = lane - 9;
                        ej = 8;
                    }
                    if (ei >= 0 && ej < nb && ei < na) {
                        float ya, yb;
                        if constexpr (kSnap) {
                            ya = (static_cast<float>(ei) - q1) * 0.25f;
                            yb = (static_cast<float>(ej) - q2) * 0.25f;
                        } else {
                            ya = (static_cast<float>(oac + a0l + ei) - q1) * 0.25f;
                            yb = (static_cast<float>(obc + b0l + ej) - q2) * 0.25f;
                        }
                        const int ca = a0l + ei, cb = b0l + ej;
                        if (ca >= 0 && ca < T && cb >= 0 && cb < T) {
                            const float ekv = es_kernel_fast(ya) * es_kernel_fast(yb);
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
                    }
                }
            }
            const int ca = a0l + ci, cb = b0l + cj;
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
        __syncthreads();

        const int G = gp.G;
        const size_t plane_size = static_cast<size_t>(G) * G;
        for (int e = threadIdx.x; e < K * T * T; e += kTileThreads) {
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

template <bool kSnap, int T, int KMAX>
void launch_tile_kernel(svdi_context* ctx, unsigned grid, const Item* items, uint32_t* totals,
                        const Records& rec, const GridParams& gp, const TileGeom& tg, float2* planes) {
    constexpr size_t bytes = TileLds<T, KMAX>::bytes;
    static_assert(bytes <= 150 * 1024, "tile LDS exceeds the CU's LDS");
    SVDI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_grid<kSnap, T, KMAX>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    k_tile_grid<kSnap, T, KMAX><<<grid, kTileThreads, bytes, ctx->stream>>>(items, totals, rec, gp, tg, planes);
    SVDI_LAUNCH_CHECK();
}

template <bool kSnap>
void launch_tile_dispatch(svdi_context* ctx, unsigned grid, const Item* items, uint32_t* totals,
                          const Records& rec, const GridParams& gp, const TileGeom& tg, float2* planes) {
    if (gp.K <= 4) launch_tile_kernel<kSnap, 32, 4>(ctx, grid, items, totals, rec, gp, tg, planes);
    else if (gp.K <= 8) launch_tile_kernel<kSnap, 32, 8>(ctx, grid, items, totals, rec, gp, tg, planes);
    else if (gp.K <= 16) launch_tile_kernel<kSnap, 16, 16>(ctx, grid, items, totals, rec, gp, tg, planes);
    else launch_tile_kernel<kSnap, 16, 32>(ctx, grid, items, totals, rec, gp, tg, planes);
}

}  // namespace

const uint32_t* launch_grid_tiled(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                                  float2* planes) {
    TileGeom tg;
    tg.T = gp.K <= 8 ? 32 : 16;
    tg.ntx = (gp.G + tg.T - 1) / tg.T;
    tg.ntiles = tg.ntx * tg.ntx;
    tg.nbins = tg.ntiles * gp.nslabs;
    tg.use_lds = tg.nbins <= kLdsBins;
    const size_t S = in.num_samples;
    const size_t max_records = 4 * S;                      // a footprint covers at most 2 x 2 tiles
    const size_t max_items = static_cast<size_t>(tg.nbins) + (max_records + kMinItem - 1) / kMinItem;
    if (max_records >= (1ull << 31)) throw UnsupportedError("too many samples for the tiled gridder");

    uint32_t* counts = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinCounts, (tg.nbins + 4) * sizeof(uint32_t)));
    uint32_t* cursor = static_cast<uint32_t*>(ctx->ws.get(Workspace::kBinCursor, tg.nbins * sizeof(uint32_t)));
    uint32_t* totals = counts + tg.nbins;
    Item* items = static_cast<Item*>(ctx->ws.get(Workspace::kItems, max_items * sizeof(Item)));
    Records rec;
    rec.q = static_cast<float4*>(ctx->ws.get(Workspace::kRecPos, max_records * sizeof(float4)));
    char* aux = static_cast<char*>(ctx->ws.get(Workspace::kRecVim, max_records * (sizeof(uint32_t) + sizeof(float))));
    rec.ab = reinterpret_cast<uint32_t*>(aux);
    rec.vy = reinterpret_cast<float*>(aux + max_records * sizeof(uint32_t));

    launch_fill(ctx, reinterpret_cast<float*>(counts), tg.nbins + 4, 0.0f);
    const size_t hist_bytes = tg.use_lds ? tg.nbins * sizeof(uint32_t) : 0;
    unsigned bin_blocks;
    if (in.from_snapshot) bin_blocks = static_cast<unsigned>((in.num_rows + kBinThreads - 1) / kBinThreads);
    else bin_blocks = static_cast<unsigned>((S + kBinThreads * kBinPerThread - 1) / (kBinThreads * kBinPerThread));
    if (in.from_snapshot) k_bin_count<true><<<bin_blocks, kBinThreads, hist_bytes, ctx->stream>>>(in, gp, tg, counts);
    else k_bin_count<false><<<bin_blocks, kBinThreads, hist_bytes, ctx->stream>>>(in, gp, tg, counts);
    SVDI_LAUNCH_CHECK();
    const int cus = std::max(1, ctx->cu_count);
    k_bin_scan<<<1, 1024, 0, ctx->stream>>>(counts, tg.nbins, cursor, items, static_cast<int>(max_items), cus, totals);
    SVDI_LAUNCH_CHECK();
    k_zero_split<<<tg.nbins, 256, 0, ctx->stream>>>(counts, totals, tg, gp.G, gp.K, planes);
    SVDI_LAUNCH_CHECK();
    if (in.from_snapshot)
        k_bin_scatter<true><<<bin_blocks, kBinThreads, hist_bytes, ctx->stream>>>(in, gp, tg, cursor, rec);
    else
        k_bin_scatter<false><<<bin_blocks, kBinThreads, hist_bytes, ctx->stream>>>(in, gp, tg, cursor, rec);
    SVDI_LAUNCH_CHECK();
    // One persistent workgroup per CU (the tile kernel's LDS admits one per CU).
    const unsigned grid = static_cast<unsigned>(std::min<size_t>(max_items, static_cast<size_t>(cus)));
    if (in.from_snapshot) launch_tile_dispatch<true>(ctx, grid, items, totals, rec, gp, tg, planes);
    else launch_tile_dispatch<false>(ctx, grid, items, totals, rec, gp, tg, planes);
    return totals;
}

}  // namespace svdi
