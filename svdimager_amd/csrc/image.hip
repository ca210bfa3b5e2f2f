// Image-plane pass (fused_interpolation, SVDIntegratedImager.cu:1295-1463; accumulation /
// scaling / coordschange / p2p / finalinterp, archive/FIPkernels.cu:463-677) and the ES grid
// correction c(t) (convolveKernel, SVDIntegratedImager.cu:1204-1216).
//
// Per pixel (x fastest; idxr = x - N/2, idyr = y - N/2):
//   t1 = cell idyr, t2 = cell idxr, t3 = -v33 + sgn(v33) sqrt(1 - (t1+v13)^2 - (t2+v23)^2),
//   sum = Re sum_slabs sum_k e^{i 2 pi t3 r3c} (k ? 2 : 1) i^k J_k(2 pi t3 h) M_{slab,k}(y, x),
//   sum *= (-1)^(idxr+idyr) / (c(|idxr|) c(|idyr|) 2.4937^2),
// then |sum| / S is splatted bilinearly at the SIN re-projection of the pixel.  t3, the
// w-phase and the Bessel sequence are FP64 (the reference evaluates them in FP32).
#include <climits>

#include "svdi_internal.hpp"

namespace svdi {
namespace {

struct GL28Table {
    double node[14];
    double weight[14];
};

GL28Table make_gl28() {
    GL28Table g;
    const int n = 28;
    for (int i = 0; i < 14; ++i) {
        double x = std::cos(kPi * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = n * (x * p1 - p0) / (x * x - 1.0);
            const double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double dp = n * (x * p1 - p0) / (x * x - 1.0);
        g.node[i] = x;
        g.weight[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    return g;
}

// c(t) = 8 / 2.4937 * sum_i ES(x_i) w_i cos(pi t 8 x_i / G), t = 0 .. N/2: the Fourier
// transform of the ES kernel by the positive half of a 28-point Gauss-Legendre rule.
__global__ void k_conv_corr(int N, int G, GL28Table gl, float* __restrict__ conv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > N / 2) return;
    const double tn = static_cast<double>(t) / static_cast<double>(G);
    double acc = 0.0;
    for (int i = 0; i < 14; ++i)
        acc += es_kernel_f64(gl.node[i]) * gl.weight[i] * cos(kPi * tn * kSupport * gl.node[i]);
    conv[t] = static_cast<float>(acc * kSupport / 2.4937047051153827);
}

// The corrected image value of pixel (x, y) (FP64 t3, w-phase and Bessel terms), or false for a
// pixel outside the unit sphere (ra < 0), which contributes nothing.
template <int KMAX>
__device__ __forceinline__ bool image_pixel(const float2* __restrict__ planes, const ImageParams& ip,
                                            const float* V, const float* __restrict__ conv, int x, int y,
                                            float& ps_out) {
    const long half = ip.N / 2;
    const long idxr = x - half, idyr = y - half;
    const double t1 = static_cast<double>(ip.cell) * static_cast<double>(idyr);
    const double t2 = static_cast<double>(ip.cell) * static_cast<double>(idxr);
    const double v13 = V[2], v23 = V[5], v33 = V[8];
    const double ra = 1.0 - (t1 + v13) * (t1 + v13) - (t2 + v23) * (t2 + v23);
    if (ra < 0.0) return false;
    // Plane element read by this pixel: linear index G*G/2 + G/2 + idyr*G + idxr of a G x G
    // plane (SVDIntegratedImager.cu:1397), located in the transformed window.  With
    // G*G/2 + G/2 = q0 G + r0 (0 <= r0 < G) and |idxr| < G, the row and column are
    // q0 + idyr and r0 + idxr with at most one carry.  The first slab's K elements are
    // requested before the FP64 work below, so that their latency overlaps it.
    const int G = ip.G;
    int lr = ip.lin_q0 + static_cast<int>(idyr), lc = ip.lin_r0 + static_cast<int>(idxr);
    if (lc < 0) {
        lc += G;
        --lr;
    } else if (lc >= G) {
        lc -= G;
        ++lr;
    }
    const long gr = lr - ip.row_lo, gc = lc - ip.col_lo;
    const size_t plane_stride = static_cast<size_t>(ip.rows) * ip.cols;
    const float2* px = planes + static_cast<size_t>(gr) * ip.cols + gc;
    float2 m[KMAX];
    auto load_slab = [&](int ls) {
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            m[k] = k < ip.K ? px[static_cast<size_t>(ls * ip.K + k) * plane_stride] : make_float2(0.0f, 0.0f);
    };
    load_slab(0);
    const double t3 = -v33 + (v33 >= 0.0 ? 1.0 : -1.0) * sqrt(ra);
    const double alpha = 2.0 * kPi * t3 * static_cast<double>(ip.slab_half);
    double jn[KMAX];
    bessel_sequence_fixed<KMAX>(alpha, jn);
    double sum = 0.0;
    for (int ls = 0; ls < ip.nslabs; ++ls) {
        const float centre = ip.archive ? ip.r3_centre
                                        : ip.r3_min + (static_cast<float>(ip.slab_start + ls) + 0.5f) * ip.slab_width;
        double sb, cb;                          // e^{i 2 pi t3 centre}
        sincospi(2.0 * t3 * static_cast<double>(centre), &sb, &cb);
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const double a = (k == 0 ? 1.0 : 2.0) * jn[k];
            double cr, ci;
            switch (k & 3) {                    // a i^k e^{i beta}
                case 0: cr = a * cb; ci = a * sb; break;
                case 1: cr = -a * sb; ci = a * cb; break;
                case 2: cr = -a * cb; ci = -a * sb; break;
                default: cr = a * sb; ci = -a * cb; break;
            }
            sum += cr * m[k].x - ci * m[k].y;   // zero for k >= K
        }
        if (ls + 1 < ip.nslabs) load_slab(ls + 1);
    }
    float ps = static_cast<float>(sum);
    if ((idxr + idyr) & 1) ps = -ps;
    ps *= 1 / (conv[labs(idxr)] * conv[labs(idyr)] * kConvCorrNorm * kConvCorrNorm);
    ps_out = ps;
    return true;
}

// Bilinear splat tile of a 16 x 16 block: the block's targets are accumulated in LDS and each
// touched cell is added to the image with one device atomic, instead of four per pixel.
constexpr int kSplatTile = 32;

template <int KMAX>
__global__ void __launch_bounds__(256)
k_image(const float2* __restrict__ planes, ImageParams ip, const float* __restrict__ vin,
        const float* __restrict__ conv, float* __restrict__ image) {
    take_device_plan(ip);
    const int x = blockIdx.x * 16 + threadIdx.x;
    const int y = blockIdx.y * 16 + threadIdx.y;
    float V[9];
    for (int i = 0; i < 9; ++i) V[i] = vin[i];
    float ps = 0.0f;
    const bool ok = x < ip.N && y < ip.N && image_pixel<KMAX>(planes, ip, V, conv, x, y, ps);
    if (ip.signed_out) {
        if (ok) image[static_cast<size_t>(y) * ip.N + x] += ps * ip.inv_num_samples;
        return;
    }
    // Every thread of the block reaches the barriers below.
    const long half = ip.N / 2;
    const float value = fabsf(ps) * ip.inv_num_samples;
    float LL = 0.0f, MM = 0.0f;
    if (ok) {
        float oi0, oi1;
        if (ip.archive) sin_reproject_archive(ip.N, ip.cell, V, x, y, &oi0, &oi1);
        else sin_reproject(ip.N, ip.cell, V, x, y, &oi0, &oi1);
        LL = oi0 - half;
        MM = oi1 - half;
    }
    const bool splat = ok && fabsf(LL) < half - 1 && fabsf(MM) < half - 1;
    const float LLf = floorf(LL), MMf = floorf(MM);
    const float LLc = ceilf(LL), MMc = ceilf(MM);
    const int lf = static_cast<int>(LLf), mf = static_cast<int>(MMf);
    const int lc = static_cast<int>(LLc), mc = static_cast<int>(MMc);
    const float w00 = (1 - LL + LLf) * (1 - MM + MMf) * value, w01 = (1 - LL + LLf) * (0 + MM - MMf) * value;
    const float w10 = (0 + LL - LLf) * (1 - MM + MMf) * value, w11 = (0 + LL - LLf) * (0 + MM - MMf) * value;
    float* d = image + half * ip.N + half;
    const long n = ip.N;
    __shared__ int s_box[4];                       // min l, max l, min m, max m of the targets
    __shared__ float s_tile[kSplatTile * kSplatTile];
    const int tid = threadIdx.y * 16 + threadIdx.x;
    if (tid == 0) {
        s_box[0] = INT_MAX;
        s_box[1] = INT_MIN;
        s_box[2] = INT_MAX;
        s_box[3] = INT_MIN;
    }
    for (int c = tid; c < kSplatTile * kSplatTile; c += 256) s_tile[c] = 0.0f;
    __syncthreads();
    {
        // Bounds per wavefront by shuffles, then one LDS atomic per wavefront and bound (the
        // same-address atomics of every thread would serialise).
        int l0 = splat ? lf : INT_MAX, l1 = splat ? lc : INT_MIN;
        int m0 = splat ? mf : INT_MAX, m1 = splat ? mc : INT_MIN;
        for (int off = 32; off > 0; off >>= 1) {
            l0 = min(l0, __shfl_xor(l0, off, 64));
            l1 = max(l1, __shfl_xor(l1, off, 64));
            m0 = min(m0, __shfl_xor(m0, off, 64));
            m1 = max(m1, __shfl_xor(m1, off, 64));
        }
        if ((tid & 63) == 0 && l1 >= l0) {
            atomicMin(&s_box[0], l0);
            atomicMax(&s_box[1], l1);
            atomicMin(&s_box[2], m0);
            atomicMax(&s_box[3], m1);
        }
    }
    __syncthreads();
    const int bx0 = s_box[0], bx1 = s_box[1], by0 = s_box[2], by1 = s_box[3];
    if (bx1 < bx0) return;                          // no target in this block
    const bool fits = bx1 - bx0 < kSplatTile && by1 - by0 < kSplatTile;
    if (!fits) {
        if (splat) {
            unsafeAtomicAdd(&d[static_cast<long>(mf) * n + lf], w00);
            unsafeAtomicAdd(&d[static_cast<long>(mc) * n + lf], w01);
            unsafeAtomicAdd(&d[static_cast<long>(mf) * n + lc], w10);
            unsafeAtomicAdd(&d[static_cast<long>(mc) * n + lc], w11);
        }
        return;
    }
    if (splat) {
        atomicAdd(&s_tile[(mf - by0) * kSplatTile + (lf - bx0)], w00);
        atomicAdd(&s_tile[(mc - by0) * kSplatTile + (lf - bx0)], w01);
        atomicAdd(&s_tile[(mf - by0) * kSplatTile + (lc - bx0)], w10);
        atomicAdd(&s_tile[(mc - by0) * kSplatTile + (lc - bx0)], w11);
    }
    __syncthreads();
    const int bw = bx1 - bx0 + 1, cells = bw * (by1 - by0 + 1);
    for (int c = tid; c < cells; c += 256) {
        const int i = c / bw, j = c - i * bw;
        const float v = s_tile[i * kSplatTile + j];
        if (v != 0.0f) unsafeAtomicAdd(&d[static_cast<long>(by0 + i) * n + (bx0 + j)], v);
    }
}

}  // namespace

void launch_conv_corr(svdi_context* ctx, size_t N, size_t G, int /*archive_variant*/, float* conv) {
    static const GL28Table gl = make_gl28();
    const int count = static_cast<int>(N / 2 + 1);
    k_conv_corr<<<(count + 255) / 256, 256, 0, ctx->stream>>>(static_cast<int>(N), static_cast<int>(G), gl, conv);
    SVDI_LAUNCH_CHECK();
}

void launch_image(svdi_context* ctx, const float2* planes, const float* vin_dev, const float* conv,
                  const ImageParams& ip_in, float* image) {
    ImageParams ip = ip_in;
    const long long G = ip.G;
    const long long base = (G * G) / 2 + G / 2;
    ip.lin_q0 = static_cast<int>(base / G);
    ip.lin_r0 = static_cast<int>(base % G);
    dim3 block(16, 16);
    dim3 grid(static_cast<unsigned>((ip.N + 15) / 16), static_cast<unsigned>((ip.N + 15) / 16));
    if (ip.K <= 4) k_image<4><<<grid, block, 0, ctx->stream>>>(planes, ip, vin_dev, conv, image);
    else if (ip.K <= 8) k_image<8><<<grid, block, 0, ctx->stream>>>(planes, ip, vin_dev, conv, image);
    else if (ip.K <= 16) k_image<16><<<grid, block, 0, ctx->stream>>>(planes, ip, vin_dev, conv, image);
    else k_image<32><<<grid, block, 0, ctx->stream>>>(planes, ip, vin_dev, conv, image);
    SVDI_LAUNCH_CHECK();
}

}  // namespace svdi
