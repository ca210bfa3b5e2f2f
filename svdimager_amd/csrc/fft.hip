// Batched 2D inverse DFT (sign +, unnormalised) of the moment planes, replacing
// cufftPlanMany + cufftExecC2C(CUFFT_INVERSE) (SVDIntegratedImager.cu:1658-1676, 1713).
//
// Two LDS-staged passes of a mixed-radix Stockham FFT (radix 8/4/2/3/5 kernels plus a direct
// butterfly for other primes up to 13).  G = ceil(3N/2) is 3 * 2^k for the usual image sizes.
// Only the central N x N of each transformed G x G plane is read by the image pass, so the row
// pass keeps the central `cols` output columns and the column pass transforms just those
// columns and keeps the central `rows` outputs.
#include "svdi_internal.hpp"

#include <algorithm>
#include <cmath>

namespace svdi {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxStages = 24;
constexpr int kMaxRadix = 16;

// Quotient q = n / d for 0 <= n < 2^27 and d <= 2^13 as (n * ceil(2^40 / d)) >> 40: the error
// term n / 2^40 stays below 1 / d, so the floor is exact without an integer divide.
struct FastDiv {
    uint64_t mul;
    __host__ __device__ uint32_t div(uint32_t n) const { return static_cast<uint32_t>((n * mul) >> 40); }
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.mul = ((1ull << 40) + d - 1) / d;
    return f;
}

struct FftStages {
    int n;
    int count;
    int radix[kMaxStages];
    FastDiv by_m[kMaxStages];      // m = n / R of the stage
    FastDiv by_ns[kMaxStages];     // ns = product of the earlier radices
};

// LDS index of element i of a transform: one float2 of padding per 32 keeps the radix-R
// Stockham writes (stride R) off shared banks.
__device__ __forceinline__ int pidx(int i) { return i + (i >> 5); }
__host__ __device__ inline size_t padded_len(size_t n) { return n + n / 32 + 1; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }   // i * a

// out[k] = sum_r v[r] e^{+2 pi i r k / R}
__device__ __forceinline__ void dft2(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

__device__ __forceinline__ void dft3(float2* v) {
    const float c = -0.5f, s = 0.86602540378443864676f;   // e^{2 pi i / 3} = c + i s
    const float2 t1 = cadd(v[1], v[2]);
    const float2 t2 = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    const float2 is = make_float2(-s * t2.y, s * t2.x);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, is);
    v[2] = csub(m, is);
}

__device__ __forceinline__ void dft4(float2* v) {
    const float2 y0 = cadd(v[0], v[2]), y1 = csub(v[0], v[2]);
    const float2 y2 = cadd(v[1], v[3]), y3 = mul_i(csub(v[1], v[3]));
    v[0] = cadd(y0, y2);
    v[2] = csub(y0, y2);
    v[1] = cadd(y1, y3);
    v[3] = csub(y1, y3);
}

__device__ __forceinline__ void dft5(float2* v) {
    // e^{2 pi i k / 5}: c1 = cos(2pi/5), c2 = cos(4pi/5), s1 = sin(2pi/5), s2 = sin(4pi/5)
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const float2 x0 = v[0];
    const float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 m2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    // i * (s1 b1 + s2 b2) and i * (s2 b1 - s1 b2)
    const float2 n1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const float2 n2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(m1, mul_i(n1));
    v[4] = csub(m1, mul_i(n1));
    v[2] = cadd(m2, mul_i(n2));
    v[3] = csub(m2, mul_i(n2));
}

__device__ __forceinline__ void dft8(float2* v) {
    const float h = 0.70710678118654752440f;
    float2 a[4], b[4];
    for (int r = 0; r < 4; ++r) {
        a[r] = cadd(v[r], v[r + 4]);
        b[r] = csub(v[r], v[r + 4]);
    }
    // b[r] *= e^{+2 pi i r / 8}
    b[1] = make_float2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y));
    b[2] = mul_i(b[2]);
    b[3] = make_float2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y));
    dft4(a);
    dft4(b);
    for (int m = 0; m < 4; ++m) {
        v[2 * m] = a[m];
        v[2 * m + 1] = b[m];
    }
}

__device__ __forceinline__ void dft_generic(float2* v, int R, const float2* __restrict__ tw, int n) {
    float2 y[kMaxRadix];
    const int step = n / R;                 // e^{2 pi i q / R} = tw[q * n / R]
    for (int k = 0; k < R; ++k) {
        float2 acc = v[0];
        for (int q = 1; q < R; ++q) acc = cadd(acc, cmul(v[q], tw[((q * k) % R) * step]));
        y[k] = acc;
    }
    for (int k = 0; k < R; ++k) v[k] = y[k];
}

template <int R>
__device__ __forceinline__ void dft_fixed(float2* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 5) dft5(v);
    else dft8(v);
}

// One Stockham stage of radix R over ntrans transforms of length n (LDS src -> LDS dst, each
// transform `pn` float2 apart, padded): butterfly j reads src[j + r m], twiddles by
// e^{+2 pi i (j mod ns) r / (ns R)} (table tw in LDS: tw[t] = e^{2 pi i t / n}), and writes
// dst[(j / ns) ns R + (j mod ns) + r ns].
template <int R>
__device__ __forceinline__ void stage_fixed(const float2* src, float2* dst, int ntrans, int n, int pn,
                                            int ns, FastDiv by_m, FastDiv by_ns, const float2* tw) {
    const int m = n / R;
    const int tstep = n / (ns * R);
    const int total = ntrans * m;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int tr = static_cast<int>(by_m.div(t));
        const int j = t - tr * m;
        const int jq = static_cast<int>(by_ns.div(j));
        const int jm = j - jq * ns;
        const float2* in = src + tr * pn;
        float2* out = dst + tr * pn;
        float2 v[R];
        v[0] = in[pidx(j)];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(in[pidx(j + r * m)], tw[jm * r * tstep]);
        dft_fixed<R>(v);
        const int d = jq * ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) out[pidx(d + r * ns)] = v[r];
    }
}

__device__ void stage_generic(const float2* src, float2* dst, int ntrans, int n, int pn, int ns, int R,
                              FastDiv by_m, FastDiv by_ns, const float2* tw) {
    const int m = n / R;
    const int tstep = n / (ns * R);
    const int total = ntrans * m;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int tr = static_cast<int>(by_m.div(t));
        const int j = t - tr * m;
        const int jq = static_cast<int>(by_ns.div(j));
        const int jm = j - jq * ns;
        const float2* in = src + tr * pn;
        float2* out = dst + tr * pn;
        float2 v[kMaxRadix];
        v[0] = in[pidx(j)];
        for (int r = 1; r < R; ++r) v[r] = cmul(in[pidx(j + r * m)], tw[jm * r * tstep]);
        dft_generic(v, R, tw, n);
        const int d = jq * ns * R + jm;
        for (int r = 0; r < R; ++r) out[pidx(d + r * ns)] = v[r];
    }
}

// In-place form of a radix-R stage: every butterfly of the stage (at most kMaxB per thread) is
// read and transformed into registers, then, after a barrier, written back to the same buffer.
// Half the LDS of the ping-pong form, so twice the transforms per workgroup.
constexpr int kMaxB = 4;
constexpr int kInplaceThreads = 512;

template <int R>
__device__ __forceinline__ void stage_inplace(float2* buf, int ntrans, int n, int pn, int ns,
                                              FastDiv by_m, FastDiv by_ns, const float2* tw) {
    const int m = n / R;
    const int tstep = n / (ns * R);
    const int total = ntrans * m;
    float2 v[kMaxB][R];
    int obase[kMaxB];
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int t = threadIdx.x + b * kInplaceThreads;
        obase[b] = -1;
        if (t < total) {
            const int tr = static_cast<int>(by_m.div(t));
            const int j = t - tr * m;
            const int jq = static_cast<int>(by_ns.div(j));
            const int jm = j - jq * ns;
            const float2* in = buf + tr * pn;
            v[b][0] = in[pidx(j)];
#pragma unroll
            for (int r = 1; r < R; ++r) v[b][r] = cmul(in[pidx(j + r * m)], tw[jm * r * tstep]);
            dft_fixed<R>(v[b]);
            obase[b] = (tr << 20) | (jq * ns * R + jm);     // tr < 16, index < 2^20
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        if (obase[b] < 0) continue;
        float2* out = buf + (obase[b] >> 20) * pn;
        const int d = obase[b] & 0xfffff;
#pragma unroll
        for (int r = 0; r < R; ++r) out[pidx(d + r * ns)] = v[b][r];
    }
    __syncthreads();
}

__device__ void stockham_inplace(float2* buf, int ntrans, int pn, const FftStages& st, const float2* tw) {
    const int n = st.n;
    int ns = 1;
    for (int stage = 0; stage < st.count; ++stage) {
        const int R = st.radix[stage];
        const FastDiv bm = st.by_m[stage], bn = st.by_ns[stage];
        switch (R) {
            case 2: stage_inplace<2>(buf, ntrans, n, pn, ns, bm, bn, tw); break;
            case 3: stage_inplace<3>(buf, ntrans, n, pn, ns, bm, bn, tw); break;
            case 4: stage_inplace<4>(buf, ntrans, n, pn, ns, bm, bn, tw); break;
            case 5: stage_inplace<5>(buf, ntrans, n, pn, ns, bm, bn, tw); break;
            default: stage_inplace<8>(buf, ntrans, n, pn, ns, bm, bn, tw); break;
        }
        ns *= R;
    }
}

// ntrans transforms of length n, `pn` float2 apart in src (LDS, padded).  Returns the buffer
// that holds the result (src or dst).
__device__ float2* stockham(float2* src, float2* dst, int ntrans, int pn, const FftStages& st,
                            const float2* tw) {
    const int n = st.n;
    int ns = 1;
    for (int stage = 0; stage < st.count; ++stage) {
        const int R = st.radix[stage];
        const FastDiv bm = st.by_m[stage], bn = st.by_ns[stage];
        switch (R) {
            case 2: stage_fixed<2>(src, dst, ntrans, n, pn, ns, bm, bn, tw); break;
            case 3: stage_fixed<3>(src, dst, ntrans, n, pn, ns, bm, bn, tw); break;
            case 4: stage_fixed<4>(src, dst, ntrans, n, pn, ns, bm, bn, tw); break;
            case 5: stage_fixed<5>(src, dst, ntrans, n, pn, ns, bm, bn, tw); break;
            case 8: stage_fixed<8>(src, dst, ntrans, n, pn, ns, bm, bn, tw); break;
            default: stage_generic(src, dst, ntrans, n, pn, ns, R, bm, bn, tw); break;
        }
        __syncthreads();
        float2* tmp = src;
        src = dst;
        dst = tmp;
        ns *= R;
    }
    return src;
}

// put(t, get(t)) for t = threadIdx.x, + blockDim.x, ... < n, with the gets of kCopyBatch
// consecutive steps issued before their puts (several global loads in flight per thread).
constexpr int kCopyBatch = 8;
template <class Get, class Put>
__device__ __forceinline__ void copy_batched(int n, Get&& get, Put&& put) {
    const int step = static_cast<int>(blockDim.x);
    int t = static_cast<int>(threadIdx.x);
    for (; t + (kCopyBatch - 1) * step < n; t += kCopyBatch * step) {
        float2 v[kCopyBatch];
#pragma unroll
        for (int j = 0; j < kCopyBatch; ++j) v[j] = get(t + j * step);
#pragma unroll
        for (int j = 0; j < kCopyBatch; ++j) put(t + j * step, v[j]);
    }
    for (; t < n; t += step) put(t, get(t));
}

// LDS of a pass: the twiddle table (n float2), then two ping-pong buffers of ntrans padded
// transforms.
__device__ __forceinline__ const float2* load_twiddles(float2* lds, const float2* __restrict__ tw, int n) {
    copy_batched(n, [&](int t) { return tw[t]; }, [&](int t, float2 v) { lds[t] = v; });
    return lds;
}

// Row pass: planes [P][G][G] -> row_out [P][G][cols] (output columns col_lo .. col_lo+cols-1).
// kInplace: radices 2/3/4/5/8 only, one LDS buffer, kInplaceThreads threads.
template <bool kInplace>
__global__ void __launch_bounds__(kInplace ? kInplaceThreads : kThreads)
k_fft_rows(const float2* __restrict__ in, float2* __restrict__ out, FftStages st,
           const float2* __restrict__ tw_g, int batch, int col_lo, int cols) {
    extern __shared__ float2 lds[];
    const int G = st.n;
    const int pn = static_cast<int>(padded_len(G));
    const float2* tw = load_twiddles(lds, tw_g, G);
    float2* buf0 = lds + G;
    float2* buf1 = kInplace ? buf0 : buf0 + batch * pn;
    const size_t plane = blockIdx.y;
    const int row0 = blockIdx.x * batch;
    const int nrows = min(batch, G - row0);
    const float2* src = in + (plane * G + row0) * static_cast<size_t>(G);
    for (int r = 0; r < nrows; ++r)
        copy_batched(G, [&](int c) { return src[static_cast<size_t>(r) * G + c]; },
                     [&](int c, float2 v) { buf0[r * pn + pidx(c)] = v; });
    __syncthreads();
    const float2* res;
    if constexpr (kInplace) {
        stockham_inplace(buf0, nrows, pn, st, tw);
        res = buf0;
    } else {
        res = stockham(buf0, buf1, nrows, pn, st, tw);
    }
    float2* dst = out + (plane * G + row0) * static_cast<size_t>(cols);
    for (int r = 0; r < nrows; ++r)
        for (int c = threadIdx.x; c < cols; c += blockDim.x)
            dst[static_cast<size_t>(r) * cols + c] = res[r * pn + pidx(col_lo + c)];
}

// One 1536-point inverse transform per wavefront (G = 24 x 64), held in registers.  With
// n = 64 n1 + l and k = k1 + 24 k2 (Cooley-Tukey, W = e^{+2 pi i / 1536}):
//   lane l:  Y[k1] = W^{l k1} sum_{n1} x[64 n1 + l] W_24^{n1 k1}   (24-point DFT as 8 x 3),
//   X[k1 + 24 k2] = sum_l Y_l[k1] W_64^{l k2},
// the 64-point DFTs across lanes done as 8 x 8 (l = 8a + b, k2 = c + 8d) through one padded LDS
// tile per wavefront: pass A takes (k1, b) and transforms over a, times W_64^{b c}; pass B takes
// (k1, c) and transforms over b, giving X[k1 + 24 c + 192 d].  load(n) reads x[n] (lane l asks
// for n = 64 n1 + l); store(k, X[k]) is called for every k, lane l holding k = l + 64 j + 192 d.
constexpr int kFftWaves = 4;                 // wavefronts (transforms) per workgroup
constexpr int kTileStride = 65;              // float2 per k1 row of the LDS tile (64 + 1 pad)

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// kTwStride: the table holds e^{2 pi i t / (1536 kTwStride)} (W_1536^m = tw[kTwStride m]).
// fft1536_wave_regs takes the lane's inputs already in registers: y[v][u] = x[64 (3u + v) + l].
template <int kTwStride = 1, class Store>
__device__ __forceinline__ void fft1536_wave_regs(float2 (&y)[3][8], float2* tile, const float2* __restrict__ tw,
                                                  int lane, Store&& store) {
    // 24-point DFT over n1 = 3u + v, k1 = s + 8t: 8-point DFTs over u, times W_24^{v s}, then
    // 3-point DFTs over v.  W_24^m = tw[64 m].
#pragma unroll
    for (int v = 0; v < 3; ++v) dft8(y[v]);
#pragma unroll
    for (int v = 1; v < 3; ++v)
#pragma unroll
        for (int s8 = 1; s8 < 8; ++s8) y[v][s8] = cmul(y[v][s8], tw[kTwStride * 64 * v * s8]);
#pragma unroll
    for (int s8 = 0; s8 < 8; ++s8) {
        float2 t3[3] = {y[0][s8], y[1][s8], y[2][s8]};
        dft3(t3);
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int k1 = s8 + 8 * t;
            // times W^{l k1} (l k1 < 1536), into the tile at [k1][l]
            tile[k1 * kTileStride + lane] = k1 == 0 ? t3[t] : cmul(t3[t], tw[kTwStride * lane * k1]);
        }
    }
    wave_lds_fence();
    // Pass A: q = lane + 64 j -> (k1, b) = (q % 24, q / 24); 8-point DFT over a, times W_64^{b c}.
    float2 w[3][8];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int q = lane + 64 * j, k1 = q % 24, b = q / 24;
#pragma unroll
        for (int a = 0; a < 8; ++a) w[j][a] = tile[k1 * kTileStride + 8 * a + b];
    }
    wave_lds_fence();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int q = lane + 64 * j, k1 = q % 24, b = q / 24;
        dft8(w[j]);
#pragma unroll
        for (int c = 1; c < 8; ++c) w[j][c] = cmul(w[j][c], tw[kTwStride * 24 * b * c]);
#pragma unroll
        for (int c = 0; c < 8; ++c) tile[k1 * kTileStride + 8 * b + c] = w[j][c];
    }
    wave_lds_fence();
    // Pass B: p = lane + 64 j -> (k1, c) = (p % 24, p / 24); 8-point DFT over b gives X[p + 192 d].
    // Every value is read before the first store (the stores may go to the tile's memory).
    float2 z[3][8];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int p = lane + 64 * j, k1 = p % 24, c = p / 24;
#pragma unroll
        for (int b = 0; b < 8; ++b) z[j][b] = tile[k1 * kTileStride + 8 * b + c];
    }
    wave_lds_fence();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int p = lane + 64 * j;
        dft8(z[j]);
#pragma unroll
        for (int d = 0; d < 8; ++d) store(p + 192 * d, z[j][d]);
    }
}

template <int kTwStride = 1, class Load, class Store>
__device__ __forceinline__ void fft1536_wave(float2* tile, const float2* __restrict__ tw, int lane,
                                             Load&& load, Store&& store) {
    float2 y[3][8];                                // y[v][u] = x[64 (3u + v) + l]
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 3; ++v) y[v][u] = load(64 * (3 * u + v) + lane);
    wave_lds_fence();                              // the tile may be where the loads came from
    fft1536_wave_regs<kTwStride>(y, tile, tw, lane, store);
}

// Row pass for G = 1536: one wavefront per row of the P x G rows; loads stride 64 per lane and
// stores of consecutive k across lanes are coalesced; only the columns col_lo .. col_lo + cols - 1
// are stored.
__global__ void __launch_bounds__(64 * kFftWaves)
k_fft_rows_1536(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
                int total_rows, int col_lo, int cols) {
    constexpr int G = 1536;
    __shared__ float2 tile_all[kFftWaves][24 * kTileStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * kFftWaves + wave;
    if (row >= total_rows) return;                 // whole wavefronts; no workgroup barrier below
    const float2* src = in + static_cast<size_t>(row) * G;
    float2* dst = out + static_cast<size_t>(row) * cols - col_lo;
    fft1536_wave(tile_all[wave], tw, lane, [&](int n) { return src[n]; },
                 [&](int k, float2 v) {
                     if (k >= col_lo && k < col_lo + cols) dst[k] = v;
                 });
}

// Workgroups are dealt to the 8 XCDs round-robin in launch order, each XCD with its own L2.
// Remapped so that consecutive logical tiles (neighbouring column strips, which share cache
// lines) run on the same XCD at about the same time.
// A bijection of [0, total): XCD x takes the contiguous tiles [x q + min(x, r), ...), q + 1 of
// them for x < r and q otherwise (q = total / 8, r = total % 8).
__device__ __forceinline__ int xcd_local_tile(int linear, int total) {
    constexpr int kXcds = 8;
    const int q = total / kXcds, r = total % kXcds;
    const int x = linear % kXcds, local = linear / kXcds;
    return x * q + min(x, r) + local;
}

// Column pass of 1536-point transforms: one workgroup per (half-plane, strip of kColStrip
// columns).  The strip is read into LDS with the columns side by side (row segments of
// kColStrip columns), each wavefront transforms one column in place (the column's LDS region is
// also its exchange tile), and the output rows in the window row_lo .. row_lo + rows - 1 are
// written back as row segments.  kSplit = 1: G = 1536, the half-planes are the planes and
// transform index k is output row k.  kSplit = 2: G = 3072 after the row pass's radix-2 step;
// half-plane q holds a (q even) or b (q odd) of plane q / 2, and k is output row 2 k + q % 2
// (twiddles W_1536^m = tw[2 m]).
constexpr int kColStrip = 4;
constexpr int kColLen = 24 * kTileStride;          // float2 per column region (>= 1536)

template <int kSplit>
__global__ void __launch_bounds__(64 * kColStrip)
k_fft_cols_1536(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
                int cols, int row_lo, int rows, int strips) {
    constexpr int H = 1536;
    __shared__ float2 strip[kColStrip * kColLen];
    const int tile = xcd_local_tile(blockIdx.x, gridDim.x);
    const int q = tile / strips;
    const int plane = q / kSplit, parity = q - plane * kSplit;
    const int c0 = (tile - q * strips) * kColStrip;
    const int nc = min(kColStrip, cols - c0);
    const float2* src = in + static_cast<size_t>(q) * H * cols + c0;
    {
        // Element (r, c) of the strip: thread t = (r mod 64) kColStrip + c loads rows r0 + 64 j,
        // all 24 loads issued before the first LDS store.
        constexpr int kLoads = H / 64;
        const int c = threadIdx.x % kColStrip, r0 = threadIdx.x / kColStrip;
        if (c < nc) {
            float2 v[kLoads];
#pragma unroll
            for (int j = 0; j < kLoads; ++j) v[j] = src[static_cast<size_t>(r0 + 64 * j) * cols + c];
#pragma unroll
            for (int j = 0; j < kLoads; ++j) strip[c * kColLen + r0 + 64 * j] = v[j];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave < nc) {
        float2* col = strip + wave * kColLen;
        fft1536_wave<kSplit>(col, tw, lane, [&](int n) { return col[n]; }, [&](int k, float2 v) { col[k] = v; });
    }
    __syncthreads();
    // Transform outputs k_lo .. k_lo + nk - 1 land in the window (output row kSplit k + parity).
    const int k_lo = max(0, (row_lo - parity + kSplit - 1) / kSplit);
    const int nk = min(H, (row_lo + rows - parity + kSplit - 1) / kSplit) - k_lo;
    float2* dst = out + static_cast<size_t>(plane) * rows * cols + c0;
#pragma unroll 8
    for (int t = threadIdx.x; t < nk * kColStrip; t += 64 * kColStrip) {
        const int k = k_lo + t / kColStrip, c = t % kColStrip;
        const int r = kSplit * k + parity - row_lo;
        if (c < nc && r >= 0 && r < rows) dst[static_cast<size_t>(r) * cols + c] = strip[c * kColLen + k];
    }
}

// G = 3072 = 2 x 1536, a radix-2 step in each direction around 1536-point transforms.
// Along a row (decimation in time): E and O are the transforms of the even and odd samples
// (twiddles W_1536^m = tw[2 m] of the 3072-point table), X[k] = E[k] + W^k O[k] and
// X[k + 1536] = E[k] - W^k O[k], W = e^{2 pi i / 3072}.  Along a column (decimation in
// frequency, done in the row pass on the row-transformed rows x_n): a_n = x_n + x_{n + 1536} and
// b_n = (x_n - x_{n + 1536}) W^n for n < 1536; the 1536-point column transforms of a and b give
// the even and odd output rows, X[2 k] = A[k] and X[2 k + 1] = B[k].

// Row pass for G = 3072: one workgroup per (plane, n < 1536) transforms rows n and n + 1536, a
// pair of wavefronts per row (even samples, odd samples, each into its own LDS tile), combines
// them and stores a_n as row n and b_n as row 1536 + n of the plane's output (the columns
// col_lo .. col_lo + cols - 1).
__global__ void __launch_bounds__(256)
k_fft_rows_3072(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
                int col_lo, int cols) {
    constexpr int G = 3072, H = 1536;
    __shared__ float2 tile_all[4][24 * kTileStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int plane = blockIdx.x / H, n = blockIdx.x - plane * H;
    const int which = wave >> 1, half = wave & 1;               // row n or n + 1536; even or odd
    const float2* src = in + (static_cast<size_t>(plane) * G + n + which * H) * G;
    float2* tile = tile_all[wave];
    fft1536_wave<2>(tile, tw, lane, [&](int m) { return src[2 * m + half]; },
                    [&](int k, float2 v) { tile[k] = v; });
    __syncthreads();
    const float2 wn = tw[n];
    float2* dst_a = out + (static_cast<size_t>(plane) * G + n) * cols - col_lo;
    float2* dst_b = dst_a + static_cast<size_t>(H) * cols;
    for (int k = threadIdx.x; k < H; k += 256) {
        const float2 wk = tw[k];
        const float2 en = tile_all[0][k], on = cmul(tile_all[1][k], wk);
        const float2 em = tile_all[2][k], om = cmul(tile_all[3][k], wk);
        const float2 xn0 = cadd(en, on), xn1 = csub(en, on);      // row n at columns k, k + 1536
        const float2 xm0 = cadd(em, om), xm1 = csub(em, om);      // row n + 1536
        if (k >= col_lo && k < col_lo + cols) {
            dst_a[k] = cadd(xn0, xm0);
            dst_b[k] = cmul(csub(xn0, xm0), wn);
        }
        const int k1 = k + H;
        if (k1 >= col_lo && k1 < col_lo + cols) {
            dst_a[k1] = cadd(xn1, xm1);
            dst_b[k1] = cmul(csub(xn1, xm1), wn);
        }
    }
}

// Row pass for G = 6144 = 4 x 1536: one workgroup per row, wavefront j transforming the samples
// x[4 m + j] (m < 1536, the 1536-point wavefront transform with a stride-4 twiddle table) into
// F_j, then a radix-4 step: with W = e^{2 pi i / 6144} and k = k0 + 1536 q (k0 < 1536, q < 4),
// X[k] = sum_j W^{j k} F_j[k0] = sum_j i^{j q} (W^{j k0} F_j[k0]).  Only the columns
// col_lo .. col_lo + cols - 1 are stored; the column pass is the general one.
__global__ void __launch_bounds__(256)
k_fft_rows_6144(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
                int col_lo, int cols) {
    constexpr int G = 6144, H = 1536;
    __shared__ float2 tile_all[4][24 * kTileStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t row = blockIdx.x;
    const float2* src = in + row * G;
    float2* tile = tile_all[wave];
    fft1536_wave<4>(tile, tw, lane, [&](int m) { return src[4 * m + wave]; },
                    [&](int k, float2 v) { tile[k] = v; });
    __syncthreads();
    float2* dst = out + row * static_cast<size_t>(cols) - col_lo;
    for (int k0 = threadIdx.x; k0 < H; k0 += 256) {
        const float2 g0 = tile_all[0][k0];
        const float2 g1 = cmul(tile_all[1][k0], tw[k0]);
        const float2 g2 = cmul(tile_all[2][k0], tw[2 * k0]);
        const float2 g3 = cmul(tile_all[3][k0], tw[3 * k0]);
        const float2 s02 = cadd(g0, g2), d02 = csub(g0, g2);
        const float2 s13 = cadd(g1, g3), d13 = csub(g1, g3);
        const float2 id13 = make_float2(-d13.y, d13.x);          // i (g1 - g3)
        const float2 x[4] = {cadd(s02, s13), cadd(d02, id13), csub(s02, s13), csub(d02, id13)};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + H * q;
            if (k >= col_lo && k < col_lo + cols) dst[k] = x[q];
        }
    }
}

// Column pass for G = 6144 = 4 x 1536: decimation in frequency along the column, then four
// 1536-point transforms.  With W = e^{2 pi i / 6144} and n < 1536,
//   y_q[n] = W^{n q} sum_m i^{m q} x[n + 1536 m],   X[4 k + q] = sum_n y_q[n] W_1536^{n k}.
// One workgroup per (plane, strip of kColStrip6 columns): the strip is read into LDS (column c's
// rows 1536 m .. 1536 m + 1535 in its region m), wavefront (c, q) forms its 24 values of y_q in
// registers from the four regions, and after a workgroup barrier transforms them in place in
// region q (the region doubles as the exchange tile); the output rows of the window are written
// back as row segments of the strip.
constexpr int kColStrip6 = 3;
constexpr int kCols6Threads = 64 * 4 * kColStrip6;
constexpr size_t kCols6Lds = sizeof(float2) * kColStrip6 * 4 * kColLen;

__global__ void __launch_bounds__(kCols6Threads)
k_fft_cols_6144(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
                int cols, int row_lo, int rows, int strips) {
    constexpr int G = 6144, H = 1536;
    constexpr int kRowsPerPass = kCols6Threads / kColStrip6;      // 256
    constexpr int kPasses = G / kRowsPerPass;                      // 24
    extern __shared__ float2 lds6[];                               // [column][region m][kColLen]
    const int tile = xcd_local_tile(blockIdx.x, gridDim.x);
    const int plane = tile / strips;
    const int c0 = (tile - plane * strips) * kColStrip6;
    const int nc = min(kColStrip6, cols - c0);
    {
        const float2* src = in + static_cast<size_t>(plane) * G * cols + c0;
        const int c = threadIdx.x % kColStrip6, r0 = threadIdx.x / kColStrip6;
        if (c < nc) {
            float2 v[kPasses];
#pragma unroll
            for (int j = 0; j < kPasses; ++j) v[j] = src[static_cast<size_t>(r0 + kRowsPerPass * j) * cols + c];
#pragma unroll
            for (int j = 0; j < kPasses; ++j) {
                const int r = r0 + kRowsPerPass * j;
                lds6[(c * 4 + r / H) * kColLen + r % H] = v[j];
            }
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = wave >> 2, q = wave & 3;
    float2 y[3][8];
    if (c < nc) {
        const float2* col = lds6 + c * 4 * kColLen;
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const int n = 64 * (3 * u + v) + lane;
                const float2 x0 = col[n], x1 = col[kColLen + n], x2 = col[2 * kColLen + n], x3 = col[3 * kColLen + n];
                float2 t;
                if (q == 0) t = cadd(cadd(x0, x2), cadd(x1, x3));
                else if (q == 1) t = cadd(csub(x0, x2), mul_i(csub(x1, x3)));
                else if (q == 2) t = csub(cadd(x0, x2), cadd(x1, x3));
                else t = csub(csub(x0, x2), mul_i(csub(x1, x3)));
                y[v][u] = q == 0 ? t : cmul(t, tw[n * q]);
            }
    }
    __syncthreads();                               // every region read before any is rewritten
    if (c < nc) {
        float2* region = lds6 + (c * 4 + q) * kColLen;
        fft1536_wave_regs<4>(y, region, tw, lane, [&](int k, float2 v) { region[k] = v; });
    }
    __syncthreads();
    float2* dst = out + static_cast<size_t>(plane) * rows * cols + c0;
    for (int t = threadIdx.x; t < rows * kColStrip6; t += kCols6Threads) {
        const int r = t / kColStrip6, cc = t - r * kColStrip6;
        const int gr = row_lo + r;                                 // output row 4 k + q
        if (cc < nc) dst[static_cast<size_t>(r) * cols + cc] = lds6[(cc * 4 + (gr & 3)) * kColLen + (gr >> 2)];
    }
}

// Column pass: row_out [P][G][cols] -> out [P][rows][cols] (output rows row_lo .. row_lo+rows-1).
// One workgroup per (plane, strip of `width` columns), strips ordered XCD-locally.
template <bool kInplace>
__global__ void __launch_bounds__(kInplace ? kInplaceThreads : kThreads)
k_fft_cols(const float2* __restrict__ in, float2* __restrict__ out, FftStages st,
           const float2* __restrict__ tw_g, int width, int cols, int row_lo, int rows, int strips) {
    extern __shared__ float2 lds[];
    const int G = st.n;
    const int pn = static_cast<int>(padded_len(G));
    const float2* tw = load_twiddles(lds, tw_g, G);
    float2* buf0 = lds + G;
    float2* buf1 = kInplace ? buf0 : buf0 + width * pn;
    const int tile = xcd_local_tile(blockIdx.x, gridDim.x);
    const size_t plane = tile / strips;
    const int c0 = (tile - static_cast<int>(plane) * strips) * width;
    const int nc = min(width, cols - c0);
    const float2* src = in + plane * G * static_cast<size_t>(cols) + c0;
    copy_batched(G * nc,
                 [&](int t) {
                     const int r = t / nc, c = t - r * nc;
                     return src[static_cast<size_t>(r) * cols + c];
                 },
                 [&](int t, float2 v) {
                     const int r = t / nc, c = t - r * nc;
                     buf0[c * pn + pidx(r)] = v;
                 });
    __syncthreads();
    const float2* res;
    if constexpr (kInplace) {
        stockham_inplace(buf0, nc, pn, st, tw);
        res = buf0;
    } else {
        res = stockham(buf0, buf1, nc, pn, st, tw);
    }
    float2* dst = out + plane * rows * static_cast<size_t>(cols) + c0;
    for (int t = threadIdx.x; t < rows * nc; t += blockDim.x) {
        const int r = t / nc, c = t - r * nc;
        dst[static_cast<size_t>(r) * cols + c] = res[c * pn + pidx(row_lo + r)];
    }
}

// Direct O(G^2) DFT for lengths with a prime factor above 13 (rows then columns).
__global__ void k_dft_rows_naive(const float2* __restrict__ in, float2* __restrict__ out, int G,
                                 const float2* __restrict__ tw, int col_lo, int cols) {
    const size_t plane = blockIdx.z;
    const int r = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int k = col_lo + c;
    const float2* src = in + (plane * G + r) * static_cast<size_t>(G);
    float2 acc = make_float2(0.f, 0.f);
    for (int n = 0; n < G; ++n)
        acc = cadd(acc, cmul(src[n], tw[(static_cast<long>(n) * k) % G]));
    out[(plane * G + r) * static_cast<size_t>(cols) + c] = acc;
}

__global__ void k_dft_cols_naive(const float2* __restrict__ in, float2* __restrict__ out, int G,
                                 const float2* __restrict__ tw, int cols, int row_lo, int rows) {
    const size_t plane = blockIdx.z;
    const int r = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int k = row_lo + r;
    const float2* src = in + plane * G * static_cast<size_t>(cols) + c;
    float2 acc = make_float2(0.f, 0.f);
    for (int n = 0; n < G; ++n)
        acc = cadd(acc, cmul(src[static_cast<size_t>(n) * cols], tw[(static_cast<long>(n) * k) % G]));
    out[(plane * rows + r) * static_cast<size_t>(cols) + c] = acc;
}

bool factorize(int n, FftStages& st) {
    st.n = n;
    st.count = 0;
    int m = n;
    const int pref[] = {8, 4, 2, 3, 5, 7, 11, 13};
    for (int r : pref) {
        while (m % r == 0 && m > 1) {
            if (st.count == kMaxStages) return false;
            st.radix[st.count++] = r;
            m /= r;
        }
    }
    int ns = 1;
    for (int i = 0; i < st.count; ++i) {
        st.by_m[i] = make_fastdiv(static_cast<uint32_t>(n / st.radix[i]));
        st.by_ns[i] = make_fastdiv(static_cast<uint32_t>(ns));
        ns *= st.radix[i];
    }
    return m == 1 && n <= 8192;
}

}  // namespace

// Twiddles e^{2 pi i t / G}, t < G, from FP64 sincospi, rounded to FP32.
__global__ void k_twiddles(float2* __restrict__ tw, int G) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= G) return;
    double sn, cs;
    sincospi(2.0 * static_cast<double>(t) / static_cast<double>(G), &sn, &cs);
    tw[t] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
}

// The twiddle table of G in the context's workspace, computed on the stream (no host copy, no
// synchronisation) when G or the table's address changed since the last call.
void fft_prepare(svdi_context* ctx, size_t G) {
    float2* d = static_cast<float2*>(ctx->ws.get(Workspace::kTwiddles, G * sizeof(float2)));
    if (ctx->twiddle_n == G && ctx->twiddle_at == d) return;
    k_twiddles<<<static_cast<unsigned>((G + 255) / 256), 256, 0, ctx->stream>>>(d, static_cast<int>(G));
    SVDI_LAUNCH_CHECK();
    ctx->twiddle_n = G;
    ctx->twiddle_at = d;
}

void launch_fft2_inverse(svdi_context* ctx, const float2* planes, size_t P, size_t G,
                         size_t row_lo, size_t rows, size_t col_lo, size_t cols,
                         float2* row_out, float2* out) {
    fft_prepare(ctx, G);
    const float2* tw = static_cast<const float2*>(ctx->ws.get(Workspace::kTwiddles, G * sizeof(float2)));
    FftStages st;
    const size_t per = 2 * padded_len(G) * sizeof(float2);     // one transform, ping-pong
    const size_t twb = G * sizeof(float2);
    const bool ok = factorize(static_cast<int>(G), st) && twb + per <= 160 * 1024;
    if (!ok) {
        dim3 gr(static_cast<unsigned>((cols + 127) / 128), static_cast<unsigned>(G), static_cast<unsigned>(P));
        k_dft_rows_naive<<<gr, 128, 0, ctx->stream>>>(planes, row_out, static_cast<int>(G), tw,
                                                      static_cast<int>(col_lo), static_cast<int>(cols));
        SVDI_LAUNCH_CHECK();
        dim3 gc(static_cast<unsigned>((cols + 127) / 128), static_cast<unsigned>(rows), static_cast<unsigned>(P));
        k_dft_cols_naive<<<gc, 128, 0, ctx->stream>>>(row_out, out, static_cast<int>(G), tw,
                                                      static_cast<int>(cols), static_cast<int>(row_lo),
                                                      static_cast<int>(rows));
        SVDI_LAUNCH_CHECK();
        return;
    }
    // In-place form when every radix has a fixed kernel and a stage's butterflies for at least
    // one transform fit kMaxB per thread: transforms per workgroup = min over stages of
    // kMaxB * threads / m, capped by 64 KB of LDS.
    bool inplace = true;
    int fit = 16;
    for (int i = 0; i < st.count; ++i) {
        const int R = st.radix[i];
        if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8) inplace = false;
        fit = std::min(fit, kMaxB * kInplaceThreads / (static_cast<int>(G) / R));
    }
    const size_t one = padded_len(G) * sizeof(float2);
    if (inplace && fit >= 1) {
        fit = static_cast<int>(std::max<size_t>(1, std::min<size_t>(fit, (64 * 1024 > twb ? (64 * 1024 - twb) / one : 1))));
    } else {
        inplace = false;
        const size_t budget = 128 * 1024;
        const size_t room = budget > twb ? (budget - twb) / per : 0;
        fit = static_cast<int>(std::max<size_t>(1, std::min<size_t>(16, room)));
    }
    const int batch = fit, width = fit;
    const size_t lds = twb + static_cast<size_t>(fit) * (inplace ? one : per);
    const int threads = inplace ? kInplaceThreads : kThreads;
    auto rows_fn = inplace ? &k_fft_rows<true> : &k_fft_rows<false>;
    auto cols_fn = inplace ? &k_fft_cols<true> : &k_fft_cols<false>;
    allow_dynamic_lds(reinterpret_cast<const void*>(rows_fn), lds);
    allow_dynamic_lds(reinterpret_cast<const void*>(cols_fn), lds);
    if (G == 6144) {
        k_fft_rows_6144<<<static_cast<unsigned>(P * G), 256, 0, ctx->stream>>>(
            planes, row_out, tw, static_cast<int>(col_lo), static_cast<int>(cols));
    } else if (G == 3072) {
        k_fft_rows_3072<<<static_cast<unsigned>(P * 1536), 256, 0, ctx->stream>>>(
            planes, row_out, tw, static_cast<int>(col_lo), static_cast<int>(cols));
    } else if (G == 1536) {
        const int total_rows = static_cast<int>(P * G);
        k_fft_rows_1536<<<static_cast<unsigned>((total_rows + kFftWaves - 1) / kFftWaves), 64 * kFftWaves, 0,
                          ctx->stream>>>(planes, row_out, tw, total_rows, static_cast<int>(col_lo),
                                         static_cast<int>(cols));
    } else {
        dim3 gr(static_cast<unsigned>((G + batch - 1) / batch), static_cast<unsigned>(P));
        rows_fn<<<gr, threads, lds, ctx->stream>>>(planes, row_out, st, tw, batch, static_cast<int>(col_lo),
                                                   static_cast<int>(cols));
    }
    SVDI_LAUNCH_CHECK();
    if (G == 6144) {
        const int strips = static_cast<int>((cols + kColStrip6 - 1) / kColStrip6);
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_fft_cols_6144), kCols6Lds);
        k_fft_cols_6144<<<static_cast<unsigned>(strips * P), kCols6Threads, kCols6Lds, ctx->stream>>>(
            row_out, out, tw, static_cast<int>(cols), static_cast<int>(row_lo), static_cast<int>(rows), strips);
    } else if (G == 3072) {
        const int strips = static_cast<int>((cols + kColStrip - 1) / kColStrip);
        k_fft_cols_1536<2><<<static_cast<unsigned>(strips * 2 * P), 64 * kColStrip, 0, ctx->stream>>>(
            row_out, out, tw, static_cast<int>(cols), static_cast<int>(row_lo), static_cast<int>(rows), strips);
    } else if (G == 1536) {
        const int strips = static_cast<int>((cols + kColStrip - 1) / kColStrip);
        k_fft_cols_1536<1><<<static_cast<unsigned>(strips * P), 64 * kColStrip, 0, ctx->stream>>>(
            row_out, out, tw, static_cast<int>(cols), static_cast<int>(row_lo), static_cast<int>(rows), strips);
    } else {
        const int strips = static_cast<int>((cols + width - 1) / width);
        const unsigned gc = static_cast<unsigned>(strips * P);
        cols_fn<<<gc, threads, lds, ctx->stream>>>(row_out, out, st, tw, width, static_cast<int>(cols),
                                                   static_cast<int>(row_lo), static_cast<int>(rows), strips);
    }
    SVDI_LAUNCH_CHECK();
}

}  // namespace svdi
