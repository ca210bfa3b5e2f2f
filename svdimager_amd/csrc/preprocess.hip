// GPU preprocess of one snapshot: the principal-axes rotation of the MFS-folded baselines.
//
// The reference (SVDIntegratedImager.cu:635-853) expands b = uvw * nu / c into a 3 x S array,
// centres it, forms the 3x3 scatter with a tall-skinny SGEMM and diagonalises it with
// cuSOLVER.  Here the scatter is formed from O(rows) moments in FP64 instead:
//   sum_j b_j b_j^T = (sum_rows uvw uvw^T) (sum_chan nu^2) / c^2,
//   mean            = (sum_rows uvw) (sum_chan nu) / (c S),
//   scatter         = sum_j b_j b_j^T - S mean mean^T,
// which is the same matrix without touching S-sized data.  The 3x3 eigenproblem, the basis
// orientation rules and the rotation of each row's uvw run on the device in FP64; per-sample
// coordinates (row rotation x nu / c) are formed later inside the gridding pass.
#include "svdi_internal.hpp"

namespace svdi {
namespace {

constexpr int kMomentBlocks = 256;
constexpr int kThreads = 256;

// Nine FP64 row moments per block: sum u, v, w and the six products.
__global__ void __launch_bounds__(kThreads)
k_row_moments(const double* __restrict__ uvw, size_t rows, double* __restrict__ partials) {
    __shared__ double red[9][kThreads];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t r = blockIdx.x * (size_t)kThreads + threadIdx.x; r < rows;
         r += (size_t)gridDim.x * kThreads) {
        const double u = uvw[3 * r], v = uvw[3 * r + 1], w = uvw[3 * r + 2];
        acc[0] += u;
        acc[1] += v;
        acc[2] += w;
        acc[3] += u * u;
        acc[4] += u * v;
        acc[5] += u * w;
        acc[6] += v * v;
        acc[7] += v * w;
        acc[8] += w * w;
    }
    for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride)
            for (int k = 0; k < 9; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x < 9) partials[blockIdx.x * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// One block: reduce the moments, build the scatter matrix, diagonalise, orient.
__global__ void __launch_bounds__(kThreads)
k_geometry_finalize(const double* __restrict__ partials, int nblocks,
                    const double* __restrict__ freq, size_t channels, size_t rows,
                    double array_center_x, DevGeometry* __restrict__ geo) {
    __shared__ double fs[4][kThreads];
    __shared__ double red[9][kThreads];
    double s1 = 0.0, s2 = 0.0, fmn = INFINITY, fmx = -INFINITY;
    for (size_t c = threadIdx.x; c < channels; c += kThreads) {
        const double f = freq[c];
        s1 += f;
        s2 += f * f;
        fmn = fmin(fmn, f);
        fmx = fmax(fmx, f);
    }
    fs[0][threadIdx.x] = s1;
    fs[1][threadIdx.x] = s2;
    fs[2][threadIdx.x] = fmn;
    fs[3][threadIdx.x] = fmx;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += kThreads)
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += partials[b * 9 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) red[i][threadIdx.x] = acc[i];
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            fs[0][threadIdx.x] += fs[0][threadIdx.x + stride];
            fs[1][threadIdx.x] += fs[1][threadIdx.x + stride];
            fs[2][threadIdx.x] = fmin(fs[2][threadIdx.x], fs[2][threadIdx.x + stride]);
            fs[3][threadIdx.x] = fmax(fs[3][threadIdx.x], fs[3][threadIdx.x + stride]);
#pragma unroll
            for (int i = 0; i < 9; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + stride];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double mom[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) mom[i] = red[i][0];
    const double c = kSpeedOfLight;
    const double S = static_cast<double>(rows) * static_cast<double>(channels);
    const double sf = fs[0][0], sf2 = fs[1][0];
    const double su[3] = {mom[0], mom[1], mom[2]};
    const double A[9] = {mom[3], mom[4], mom[5], mom[4], mom[6], mom[7], mom[5], mom[7], mom[8]};
    double mean[3];
    for (int i = 0; i < 3; ++i) mean[i] = su[i] * sf / (c * S);
    double M[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i * 3 + j] = A[i * 3 + j] * (sf2 / (c * c)) - S * mean[i] * mean[j];
    double ev[3], evec[9], R[9];
    eigen_sym3(M, ev, evec);
    oriented_basis(evec, array_center_x, R);
    for (int i = 0; i < 9; ++i) {
        geo->R[i] = R[i];
        geo->vin[i] = static_cast<float>(R[i]);
    }
    for (int i = 0; i < 3; ++i) {
        geo->evals[i] = ev[i];
        geo->mean[i] = mean[i];
    }
    geo->fmin = fs[2][0];
    geo->fmax = fs[3][0];
    geo->det = static_cast<float>(fabs(R[0] * R[4] - R[1] * R[3]));
}

// rot[row] = R uvw[row] (metres, FP64) and per-block r3 range over all channels:
// r3(row, chan) = rot3 * (nu_chan / c) is monotone in nu, so its extremes sit at fmin / fmax.
__global__ void __launch_bounds__(kThreads)
k_rotate_rows(const double* __restrict__ uvw, size_t rows, const DevGeometry* __restrict__ geo,
              double* __restrict__ rot, float* __restrict__ r3_partials) {
    __shared__ float rmin[kThreads], rmax[kThreads];
    const size_t r = blockIdx.x * (size_t)kThreads + threadIdx.x;
    float lo = INFINITY, hi = -INFINITY;
    if (r < rows) {
        const double u = uvw[3 * r], v = uvw[3 * r + 1], w = uvw[3 * r + 2];
        double out[3];
        for (int k = 0; k < 3; ++k) out[k] = geo->R[3 * k] * u + geo->R[3 * k + 1] * v + geo->R[3 * k + 2] * w;
        rot[3 * r] = out[0];
        rot[3 * r + 1] = out[1];
        rot[3 * r + 2] = out[2];
        const double a = out[2] * (geo->fmin / kSpeedOfLight);
        const double b = out[2] * (geo->fmax / kSpeedOfLight);
        lo = static_cast<float>(fmin(a, b));
        hi = static_cast<float>(fmax(a, b));
    }
    rmin[threadIdx.x] = lo;
    rmax[threadIdx.x] = hi;
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            rmin[threadIdx.x] = fminf(rmin[threadIdx.x], rmin[threadIdx.x + stride]);
            rmax[threadIdx.x] = fmaxf(rmax[threadIdx.x], rmax[threadIdx.x + stride]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        r3_partials[2 * blockIdx.x] = rmin[0];
        r3_partials[2 * blockIdx.x + 1] = rmax[0];
    }
}

// Bin (S x 3 FP32) and the collapsed visibilities, as preprocess_measurement_set_gpu outputs
// them (project_bin_kernel SVDIntegratedImager.cu:509-536, collapse_visibility_kernel :588-617).
__global__ void __launch_bounds__(kThreads)
k_materialize(const double* __restrict__ uvw, const double* __restrict__ freq, size_t rows,
              const DevGeometry* __restrict__ geo, svdi_snapshot snap, float* __restrict__ bin,
              float* __restrict__ vis_re, float* __restrict__ vis_im) {
    const size_t row = blockIdx.x * (size_t)kThreads + threadIdx.x;
    const size_t chan = blockIdx.y;
    if (row >= rows) return;
    const size_t idx = row + chan * rows;
    const double sc = freq[chan] / kSpeedOfLight;
    const double u = uvw[3 * row] * sc, v = uvw[3 * row + 1] * sc, w = uvw[3 * row + 2] * sc;
    for (int k = 0; k < 3; ++k)
        bin[3 * idx + k] = static_cast<float>(geo->R[3 * k] * u + geo->R[3 * k + 1] * v + geo->R[3 * k + 2] * w);
    const float w0 = snap.flag0[idx] ? 0.0f : (snap.weight0 ? snap.weight0[idx] : 1.0f);
    const float w3 = snap.flag3[idx] ? 0.0f : (snap.weight3 ? snap.weight3[idx] : 1.0f);
    const float ws = w0 + w3;
    if (ws > 0.0f) {
        vis_re[idx] = (snap.vis0_re[idx] * w0 + snap.vis3_re[idx] * w3) / ws;
        vis_im[idx] = (snap.vis0_im[idx] * w0 + snap.vis3_im[idx] * w3) / ws;
    } else {
        vis_re[idx] = 0.0f;
        vis_im[idx] = 0.0f;
    }
}

__global__ void k_copy_vin(const DevGeometry* __restrict__ geo, float* __restrict__ vin) {
    if (threadIdx.x < 9) vin[threadIdx.x] = geo->vin[threadIdx.x];
}

}  // namespace

void launch_preprocess_geometry(svdi_context* ctx, const svdi_snapshot& snap, DevGeometry* geo,
                                double* rot, float* r3_partials, int* r3_blocks) {
    double* partials = static_cast<double*>(ctx->ws.get(Workspace::kPartials, kMomentBlocks * 9 * sizeof(double)));
    k_row_moments<<<kMomentBlocks, kThreads, 0, ctx->stream>>>(snap.uvw, snap.num_rows, partials);
    SVDI_LAUNCH_CHECK();
    k_geometry_finalize<<<1, kThreads, 0, ctx->stream>>>(partials, kMomentBlocks, snap.freq_hz,
                                                        snap.num_channels, snap.num_rows,
                                                        snap.array_center_x, geo);
    SVDI_LAUNCH_CHECK();
    const int blocks = static_cast<int>((snap.num_rows + kThreads - 1) / kThreads);
    if (rot != nullptr) {
        k_rotate_rows<<<blocks, kThreads, 0, ctx->stream>>>(snap.uvw, snap.num_rows, geo, rot, r3_partials);
        SVDI_LAUNCH_CHECK();
    }
    if (r3_blocks) *r3_blocks = blocks;
}

void launch_materialize(svdi_context* ctx, const svdi_snapshot& snap, const DevGeometry* geo,
                        float* bin, float* vin, float* vis_re, float* vis_im) {
    dim3 grid(static_cast<unsigned>((snap.num_rows + kThreads - 1) / kThreads),
              static_cast<unsigned>(snap.num_channels));
    k_materialize<<<grid, kThreads, 0, ctx->stream>>>(snap.uvw, snap.freq_hz, snap.num_rows, geo,
                                                     snap, bin, vis_re, vis_im);
    SVDI_LAUNCH_CHECK();
    k_copy_vin<<<1, 64, 0, ctx->stream>>>(geo, vin);
    SVDI_LAUNCH_CHECK();
}

}  // namespace svdi
