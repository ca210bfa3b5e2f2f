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

struct FftStages {
    int n;
    int count;
    int radix[kMaxStages];
};

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

// One Stockham stage of radix R over ntrans transforms of length n (LDS src -> LDS dst):
// butterfly j reads src[j + r m], twiddles by e^{+2 pi i (j mod ns) r / (ns R)}, and writes
// dst[(j / ns) ns R + (j mod ns) + r ns].
template <int R>
__device__ __forceinline__ void stage_fixed(const float2* src, float2* dst, int ntrans, int n, int ns,
                                            const float2* __restrict__ tw) {
    const int m = n / R;
    const int tstep = n / (ns * R);
    const int total = ntrans * m;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int tr = t / m;
        const int j = t - tr * m;
        const float2* in = src + tr * n;
        float2* out = dst + tr * n;
        const int jm = j % ns;
        float2 v[R];
        v[0] = in[j];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(in[j + r * m], tw[jm * r * tstep]);
        dft_fixed<R>(v);
        const int d = (j / ns) * ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) out[d + r * ns] = v[r];
    }
}

__device__ void stage_generic(const float2* src, float2* dst, int ntrans, int n, int ns, int R,
                              const float2* __restrict__ tw) {
    const int m = n / R;
    const int tstep = n / (ns * R);
    const int total = ntrans * m;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int tr = t / m;
        const int j = t - tr * m;
        const float2* in = src + tr * n;
        float2* out = dst + tr * n;
        const int jm = j % ns;
        float2 v[kMaxRadix];
        v[0] = in[j];
        for (int r = 1; r < R; ++r) v[r] = cmul(in[j + r * m], tw[jm * r * tstep]);
        dft_generic(v, R, tw, n);
        const int d = (j / ns) * ns * R + jm;
        for (int r = 0; r < R; ++r) out[d + r * ns] = v[r];
    }
}

// ntrans transforms of length n, stored back to back in src (LDS).  Returns the buffer that
// holds the result (src or dst).
__device__ float2* stockham(float2* src, float2* dst, int ntrans, const FftStages& st,
                            const float2* __restrict__ tw) {
    const int n = st.n;
    int ns = 1;
    for (int stage = 0; stage < st.count; ++stage) {
        const int R = st.radix[stage];
        switch (R) {
            case 2: stage_fixed<2>(src, dst, ntrans, n, ns, tw); break;
            case 3: stage_fixed<3>(src, dst, ntrans, n, ns, tw); break;
            case 4: stage_fixed<4>(src, dst, ntrans, n, ns, tw); break;
            case 5: stage_fixed<5>(src, dst, ntrans, n, ns, tw); break;
            case 8: stage_fixed<8>(src, dst, ntrans, n, ns, tw); break;
            default: stage_generic(src, dst, ntrans, n, ns, R, tw); break;
        }
        __syncthreads();
        float2* tmp = src;
        src = dst;
        dst = tmp;
        ns *= R;
    }
    return src;
}

// Row pass: planes [P][G][G] -> row_out [P][G][cols] (output columns col_lo .. col_lo+cols-1).
__global__ void __launch_bounds__(kThreads)
k_fft_rows(const float2* __restrict__ in, float2* __restrict__ out, FftStages st,
           const float2* __restrict__ tw, int batch, int col_lo, int cols) {
    extern __shared__ float2 lds[];
    const int G = st.n;
    const size_t plane = blockIdx.y;
    const int row0 = blockIdx.x * batch;
    const int nrows = min(batch, G - row0);
    const float2* src = in + (plane * G + row0) * static_cast<size_t>(G);
    for (int t = threadIdx.x; t < nrows * G; t += blockDim.x) lds[t] = src[t];
    __syncthreads();
    const float2* res = stockham(lds, lds + batch * G, nrows, st, tw);
    float2* dst = out + (plane * G + row0) * static_cast<size_t>(cols);
    for (int t = threadIdx.x; t < nrows * cols; t += blockDim.x) {
        const int r = t / cols, c = t - r * cols;
        dst[static_cast<size_t>(r) * cols + c] = res[r * G + col_lo + c];
    }
}

// Column pass: row_out [P][G][cols] -> out [P][rows][cols] (output rows row_lo .. row_lo+rows-1).
__global__ void __launch_bounds__(kThreads)
k_fft_cols(const float2* __restrict__ in, float2* __restrict__ out, FftStages st,
           const float2* __restrict__ tw, int width, int cols, int row_lo, int rows) {
    extern __shared__ float2 lds[];
    const int G = st.n;
    const size_t plane = blockIdx.y;
    const int c0 = blockIdx.x * width;
    const int nc = min(width, cols - c0);
    const float2* src = in + plane * G * static_cast<size_t>(cols) + c0;
    for (int t = threadIdx.x; t < G * nc; t += blockDim.x) {
        const int r = t / nc, c = t - r * nc;
        lds[c * G + r] = src[static_cast<size_t>(r) * cols + c];
    }
    __syncthreads();
    const float2* res = stockham(lds, lds + width * G, nc, st, tw);
    float2* dst = out + plane * rows * static_cast<size_t>(cols) + c0;
    for (int t = threadIdx.x; t < rows * nc; t += blockDim.x) {
        const int r = t / nc, c = t - r * nc;
        dst[static_cast<size_t>(r) * cols + c] = res[c * G + row_lo + r];
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
    return m == 1;
}

}  // namespace

void fft_prepare(svdi_context* ctx, size_t G) {
    if (ctx->twiddle_n == G) return;
    std::vector<float2> tw(G);
    for (size_t t = 0; t < G; ++t) {
        const double a = 2.0 * kPi * static_cast<double>(t) / static_cast<double>(G);
        tw[t] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    void* d = ctx->ws.get(Workspace::kTwiddles, G * sizeof(float2));
    SVDI_HIP(hipMemcpyAsync(d, tw.data(), G * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
    SVDI_HIP(hipStreamSynchronize(ctx->stream));
    ctx->twiddle_n = G;
}

void launch_fft2_inverse(svdi_context* ctx, const float2* planes, size_t P, size_t G,
                         size_t row_lo, size_t rows, size_t col_lo, size_t cols,
                         float2* row_out, float2* out) {
    fft_prepare(ctx, G);
    const float2* tw = static_cast<const float2*>(ctx->ws.get(Workspace::kTwiddles, G * sizeof(float2)));
    FftStages st;
    const bool ok = factorize(static_cast<int>(G), st) && G * 2 * sizeof(float2) <= 160 * 1024;
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
    const size_t per = 2 * G * sizeof(float2);
    const int batch = static_cast<int>(std::max<size_t>(1, std::min<size_t>(16, (64 * 1024) / per)));
    const int width = static_cast<int>(std::max<size_t>(1, std::min<size_t>(16, (96 * 1024) / per)));
    SVDI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_rows),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(batch * per)));
    SVDI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_cols),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(width * per)));
    dim3 gr(static_cast<unsigned>((G + batch - 1) / batch), static_cast<unsigned>(P));
    k_fft_rows<<<gr, kThreads, batch * per, ctx->stream>>>(planes, row_out, st, tw, batch,
                                                           static_cast<int>(col_lo), static_cast<int>(cols));
    SVDI_LAUNCH_CHECK();
    dim3 gc(static_cast<unsigned>((cols + width - 1) / width), static_cast<unsigned>(P));
    k_fft_cols<<<gc, kThreads, width * per, ctx->stream>>>(row_out, out, st, tw, width,
                                                           static_cast<int>(cols), static_cast<int>(row_lo),
                                                           static_cast<int>(rows));
    SVDI_LAUNCH_CHECK();
}

}  // namespace svdi
