// ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker (or the timed CPU baseline),
// never as the product path.
//
// CPU restatement of the egbdfX/SVDimager imaging path.  "ref32" functions follow the
// reference's FP32 kernels operation by operation, with device atomics replaced by a
// deterministic serial order.  "f64" helpers are FP64 checkers (DFT, FFT, preprocess).
//
// Reference files: SVDIntegratedImager.cu (cited "SVDI.cu") and archive/FIPkernels.cu
// (cited "FIPk.cu").  Every function cites the lines it restates.
//
// Pinning (what checks each function independently of this restatement) is listed in
// oracle/README.md and exercised by tests/test_oracle_*.py.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;   // the reference's M_PI is this double literal
constexpr float kSpeedOfLightF = 299792458.0f;    // SVDI.cu:382
constexpr float kBeta = 15.3704324328f;           // SVDI.cu:1235 (ES shape parameter)
constexpr int kSupport = 8;                       // SVDI.cu:1234
constexpr int kChebMax = 32;                      // SVDI.cu:35
constexpr float kChebTargetError = 1.0e-2f;       // SVDI.cu:36

// ---------------------------------------------------------------------------------------
// Gauss-Legendre 28-point rule, positive half, descending nodes (SVDI.cu:867-876 lists the
// same 14 nodes/weights rounded to 8 digits).  Computed here from P_28 by Newton iteration.
// ---------------------------------------------------------------------------------------
struct GL28 {
    double node[14];
    double weight[14];
    GL28() {
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
            node[i] = x;
            weight[i] = 2.0 / ((1.0 - x * x) * dp * dp);
        }
    }
};
const GL28& gl28() {
    static GL28 g;
    return g;
}

double es_f64(double x) {
    const double xx = x * x;
    return xx > 1.0 ? 0.0 : std::exp(15.3704324328 * (std::sqrt(1.0 - xx) - 1.0));
}

// exp_semicircle, SVDI.cu:884-888 / FIPk.cu:60-64 (FP32).
float es_f32(float beta, float x) {
    const float xx = x * x;
    return (xx > 1.0f) ? 0.0f : std::exp(beta * (std::sqrt(1.0f - xx) - 1.0f));
}

// chebyshev_sequence, SVDI.cu:890-899.
void cheb_seq_f32(float s, int n, float* v) {
    v[0] = 1.0f;
    if (n == 1) return;
    v[1] = s;
    for (int k = 2; k < n; ++k) v[k] = 2.0f * s * v[k - 1] - v[k - 2];
}

// bessel_sequence, SVDI.cu:901-920 (FP32 power series with relative early exit).
void bessel_seq_f32(float alpha, int max_order, float* j) {
    const float half_alpha = 0.5f * alpha;
    const float q = half_alpha * half_alpha;
    for (int n = 0; n <= max_order; ++n) {
        float term = 1.0f;
        for (int m = 1; m <= n; ++m) term *= half_alpha / static_cast<float>(m);
        float sum = term;
        for (int k = 1; k < 32; ++k) {
            term *= -q / (static_cast<float>(k) * static_cast<float>(n + k));
            sum += term;
            if (std::fabs(term) < 1.0e-7f * (1.0f + std::fabs(sum))) break;
        }
        j[n] = sum;
    }
}

// Jacobi-Anger coefficient of term k: (k ? 2 : 1) * i^k * J_k, rotated by e^{i beta}.
// SVDI.cu:1405-1434 / FIPk.cu:495-521.
inline void ja_coeff(int term, float jk, float cb, float sb, float* cr, float* ci) {
    if (term == 0) {
        *cr = cb * jk;
        *ci = sb * jk;
        return;
    }
    const float a = 2.0f * jk;
    switch (term & 3) {
        case 0: *cr = a * cb;  *ci = a * sb;  break;
        case 1: *cr = -a * sb; *ci = a * cb;  break;
        case 2: *cr = -a * cb; *ci = -a * sb; break;
        default: *cr = a * sb; *ci = -a * cb; break;
    }
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi in FP64.  Eigenvalues ascending,
// eigenvectors in columns (column-major), as cusolverDnSsyevd returns them (SVDI.cu:736-760).
void jacobi3(const double a_in[9], double evals[3], double evecs_colmajor[9]) {
    double a[3][3], v[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            a[r][c] = 0.5 * (a_in[r * 3 + c] + a_in[c * 3 + r]);
            v[r][c] = (r == c) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        const double diag = std::fabs(a[0][0]) + std::fabs(a[1][1]) + std::fabs(a[2][2]);
        if (off <= 1e-300 || off <= 1e-17 * diag) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0);
                const double s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int x, int y) { return a[x][x] < a[y][y]; });
    for (int j = 0; j < 3; ++j) {
        evals[j] = a[idx[j]][idx[j]];
        for (int r = 0; r < 3; ++r) evecs_colmajor[j * 3 + r] = v[r][idx[j]];
    }
}

// build_basis_kernel + orient_vin_from_array_center_kernel, SVDI.cu:490-507 and :538-586.
template <typename T>
void basis_from_evecs(const T evecs_colmajor[9], T array_center_x, T basis[9]) {
    for (int axis = 0; axis < 3; ++axis) {
        basis[0 * 3 + axis] = evecs_colmajor[axis + 2 * 3];
        basis[1 * 3 + axis] = evecs_colmajor[axis + 1 * 3];
        basis[2 * 3 + axis] = evecs_colmajor[axis + 0 * 3];
    }
    if (basis[0] < T(0))
        for (int a = 0; a < 3; ++a) basis[a] = -basis[a];
    if (array_center_x != T(0) && basis[4] * array_center_x < T(0))
        for (int a = 0; a < 3; ++a) basis[3 + a] = -basis[3 + a];
    const T r00 = basis[0], r01 = basis[1], r02 = basis[2];
    const T r10 = basis[3], r11 = basis[4], r12 = basis[5];
    T r20 = r01 * r12 - r02 * r11;
    T r21 = r02 * r10 - r00 * r12;
    T r22 = r00 * r11 - r01 * r10;
    const T norm = std::sqrt(r20 * r20 + r21 * r21 + r22 * r22);
    if (norm > T(0)) {
        r20 /= norm;
        r21 /= norm;
        r22 /= norm;
    }
    basis[6] = r20;
    basis[7] = r21;
    basis[8] = r22;
}

// FP64 mixed-radix DFT (any length), sign +, unnormalised.  Iterative Stockham over the prime
// factors of n with a direct radix-p butterfly.
void fft1d_inverse_f64(std::complex<double>* data, size_t n, std::complex<double>* work) {
    if (n <= 1) return;
    std::vector<size_t> radices;
    size_t m = n;
    for (size_t p = 2; p * p <= m; ++p)
        while (m % p == 0) {
            radices.push_back(p);
            m /= p;
        }
    if (m > 1) radices.push_back(m);
    std::complex<double>* in = data;
    std::complex<double>* out = work;
    size_t ns = 1;
    std::vector<std::complex<double>> v, y, root, tw;
    for (size_t r : radices) {
        v.resize(r);
        y.resize(r);
        root.resize(r);
        for (size_t q = 0; q < r; ++q)
            root[q] = std::polar(1.0, 2.0 * kPi * static_cast<double>(q) / static_cast<double>(r));
        // The stage's twiddles e^{i ang(jm) q}, ang(jm) = 2 pi jm / (ns r), once per (jm, q).
        tw.resize(ns * r);
        for (size_t jm = 0; jm < ns; ++jm) {
            const double ang = 2.0 * kPi * static_cast<double>(jm) / static_cast<double>(ns * r);
            for (size_t q = 0; q < r; ++q) tw[jm * r + q] = std::polar(1.0, ang * static_cast<double>(q));
        }
        const size_t stride = n / r;
        for (size_t j = 0; j < stride; ++j) {
            const size_t jm = j % ns;
            for (size_t q = 0; q < r; ++q)
                v[q] = in[j + q * stride] * tw[jm * r + q];
            for (size_t k = 0; k < r; ++k) {
                std::complex<double> acc(0.0, 0.0);
                for (size_t q = 0; q < r; ++q) acc += v[q] * root[(q * k) % r];
                y[k] = acc;
            }
            const size_t dst = (j / ns) * ns * r + jm;
            for (size_t k = 0; k < r; ++k) out[dst + k * ns] = y[k];
        }
        ns *= r;
        std::swap(in, out);
    }
    if (in != data) std::memcpy(data, in, n * sizeof(std::complex<double>));
}

// Reprojection target of one pixel, fused form of SVDI.cu:1338-1379.
void reproject_svdi(size_t di, float dc_rad, const float V[9], size_t idx, size_t idy,
                    float* oi0_out, float* oi1_out) {
    const float di2 = di * 0.5f;
    const float idxf = idx - di2;
    const float idyf = idy - di2;
    const float r0 = static_cast<float>(180.0 / kPi);
    const float dc = static_cast<float>(dc_rad / kPi * 180);
    const float V00 = V[0], V01 = V[1], V10 = V[3], V11 = V[4];
    const float V20 = V[6], V21 = V[7], V22 = V[8];
    const float xi = V20 / V22;
    const float eta = V21 / V22;
    const float p1 = (-V00 * idxf + V10 * idyf) / std::fabs(V22) + di2;
    const float p2 = (-V01 * idxf + V11 * idyf) / std::fabs(V22) + di2;
    float x = -dc * (p1 - (di2 + 1.0f));
    float y = dc * (p2 - (di2 + 1.0f));
    const float h = std::hypot(x, y);
    const float hr0 = h / r0;
    float r, w, z;
    if (h != 0.0f) {
        x /= h;
        y /= -h;
    } else {
        x = 0.0f;
        y = 1.0f;
    }
    float oi0, oi1;
    if (h <= r0) {
        r = h;
        if (h < r0 * std::sqrt(0.5f)) {
            z = 1.0f - std::sqrt(1.0f - hr0 * hr0);
        } else {
            z = hr0 * hr0 / (1.0f + std::sqrt(1.0f - hr0 * hr0));
        }
        w = xi * xi + eta * eta;
        if (w == 0.0f) {
            x = r * x;
            y = -r * y;
        } else {
            x = r * x + z * r0 * xi;
            y = -r * y + z * r0 * eta;
        }
        oi0 = -x / dc + di2 + 1.0f;
        oi1 = y / dc + di2 + 1.0f;
    } else {
        oi0 = p1;
        oi1 = p2;
    }
    *oi0_out = oi0;
    *oi1_out = oi1;
}

// WCS trig helpers of the archive (FIPk.cu:50-166), FP32.
long floor_dev(float v) { return v < 0 ? -static_cast<long>(std::ceil(-v)) : static_cast<long>(std::floor(v)); }
long ceil_dev(float v) { return v < 0 ? -static_cast<long>(std::floor(-v)) : static_cast<long>(std::ceil(v)); }
void sincosd(float angle, float& s, float& c) {
    if (std::fmod(angle, 90.0f) == 0) {
        const int i = static_cast<int>(std::fabs(static_cast<float>(floor_dev(angle / 90.0f + 0.5f)))) % 4;
        switch (i) {
            case 0: s = 0.0f; c = 1.0f; break;
            case 1: s = (angle > 0.0f) ? 1.0f : -1.0f; c = 0.0f; break;
            case 2: s = 0.0f; c = -1.0f; break;
            default: s = (angle > 0.0f) ? -1.0f : 1.0f; c = 0.0f; break;
        }
    } else {
        s = std::sin(static_cast<float>(angle * kPi / 180.0f));
        c = std::cos(static_cast<float>(angle * kPi / 180.0f));
    }
}
float sind(float angle) {
    if (std::fmod(angle, 90.0f) == 0) {
        const int i = static_cast<int>(std::fabs(static_cast<float>(floor_dev(angle / 90.0f - 0.5f)))) % 4;
        switch (i) {
            case 0: return 1.0f;
            case 1: return 0.0f;
            case 2: return -1.0f;
            default: return 0.0f;
        }
    }
    return std::sin(static_cast<float>(angle * kPi / 180.0f));
}
float cosd(float angle) {
    if (std::fmod(angle, 90.0f) == 0) {
        const int i = static_cast<int>(std::fabs(static_cast<float>(floor_dev(angle / 90.0f + 0.5f)))) % 4;
        switch (i) {
            case 0: return 1.0f;
            case 1: return 0.0f;
            case 2: return -1.0f;
            default: return 0.0f;
        }
    }
    return std::cos(static_cast<float>(angle * kPi / 180.0f));
}
float atan2d(float y, float x) {
    if (y == 0.0f) return (x >= 0.0f) ? 0.0f : 180.0f;
    if (x == 0.0f) return (y > 0.0f) ? 90.0f : -90.0f;
    return static_cast<float>(std::atan2(y, x) * 180.0f / kPi);
}
float acosd(float v) {
    if (v >= 1.0f && v - 1.0f < 1e-10) return 0.0f;
    if (v == 0.0f) return 90.0f;
    if (v <= -1.0f && v + 1.0f > -1e-10) return 180.0f;
    return static_cast<float>(std::acos(v) * 180.0f / kPi);
}
float asind(float v) {
    if (v <= -1.0f && v + 1.0f > -1e-10) return -90.0f;
    if (v == 0.0f) return 0.0f;
    if (v >= 1.0f && v - 1.0f < 1e-10) return 90.0f;
    return static_cast<float>(std::asin(v) * 180.0f / kPi);
}

// coordschange + p2p of the archive for one pixel, FIPk.cu:537-640.
void reproject_archive(size_t di, float dc_in, const float V[9], size_t idx, size_t idy,
                       float* out0, float* out1) {
    const size_t half = di / 2;
    float p1 = (-V[0] * (static_cast<float>(idx) - static_cast<float>(half)) +
                V[3] * (static_cast<float>(idy) - static_cast<float>(half))) / std::fabs(V[8]) +
               static_cast<float>(half);
    float p2 = (-V[1] * (static_cast<float>(idx) - static_cast<float>(half)) +
                V[4] * (static_cast<float>(idy) - static_cast<float>(half))) / std::fabs(V[8]) +
               static_cast<float>(half);
    *out0 = p1;
    *out1 = p2;
    const float xi = V[6] / V[8];
    const float eta = V[7] / V[8];
    const float dc = static_cast<float>(dc_in / kPi * 180);
    float x = -dc * (p1 - (static_cast<float>(di) / 2.0f + 1.0f));
    float y = dc * (p2 - (static_cast<float>(di) / 2.0f + 1.0f));
    float r0 = static_cast<float>(180.0f / kPi);
    const float x0 = x / r0;
    const float y0 = y / r0;
    const float r2 = x0 * x0 + y0 * y0;
    const float phi = (r2 != 0.0f) ? atan2d(x0, -y0) : 0.0f;
    float theta;
    if (r2 < 0.5f) {
        theta = acosd(std::sqrt(r2));
    } else if (r2 <= 1.0f) {
        theta = asind(std::sqrt(1.0f - r2));
    } else {
        return;
    }
    float sinphi, cosphi;
    sincosd(phi, sinphi, cosphi);
    x = sinphi;
    y = cosphi;
    const float t = static_cast<float>((90.0f - std::fabs(theta)) / 180.0f * kPi);
    float z, costhe;
    if (t < 1.0e-5f) {
        z = (theta > 0.0f) ? t * t / 2.0f : 2.0f - t * t / 2.0f;
        costhe = t;
    } else {
        z = 1.0f - sind(theta);
        costhe = cosd(theta);
    }
    r0 = static_cast<float>(180.0f / kPi);
    const float r = r0 * costhe;
    const float w = xi * xi + eta * eta;
    if (w == 0.0f) {
        x = r * x;
        y = -r * y;
    } else {
        z = z * r0;
        x = r * x + xi * z;
        y = -r * y + eta * z;
    }
    *out0 = -1.0f / dc * x + static_cast<float>(di) / 2.0f + 1.0f;
    *out1 = 1.0f / dc * y + static_cast<float>(di) / 2.0f + 1.0f;
}

// Bilinear splat of one value, SVDI.cu:1447-1461 / FIPk.cu:642-677.  dirty is the full image.
void splat(float* dirty, size_t di, float oi0, float oi1, float value) {
    const long half = static_cast<long>(di / 2);
    const float LL = oi0 - half;
    const float MM = oi1 - half;
    if (std::fabs(LL) < half - 1 && std::fabs(MM) < half - 1) {
        const float LLf = std::floor(LL), MMf = std::floor(MM);
        const float LLc = std::ceil(LL), MMc = std::ceil(MM);
        float* d = dirty + half * static_cast<long>(di) + half;
        d[static_cast<long>(MMf) * static_cast<long>(di) + static_cast<long>(LLf)] += (1 - LL + LLf) * (1 - MM + MMf) * value;
        d[static_cast<long>(MMc) * static_cast<long>(di) + static_cast<long>(LLf)] += (1 - LL + LLf) * (0 + MM - MMf) * value;
        d[static_cast<long>(MMf) * static_cast<long>(di) + static_cast<long>(LLc)] += (0 + LL - LLf) * (1 - MM + MMf) * value;
        d[static_cast<long>(MMc) * static_cast<long>(di) + static_cast<long>(LLc)] += (0 + LL - LLf) * (0 + MM - MMf) * value;
    }
}

int thread_count() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // namespace

extern "C" {

// ------------------------------- tables ------------------------------------------------
// GL-28 positive half (descending) and ES at the nodes, FP64.
void orc_gl28(double* nodes, double* weights, double* es_at_nodes) {
    const GL28& g = gl28();
    for (int i = 0; i < 14; ++i) {
        nodes[i] = g.node[i];
        weights[i] = g.weight[i];
        es_at_nodes[i] = es_f64(g.node[i]);
    }
}

double orc_conv_corr_norm_factor() {
    const GL28& g = gl28();
    double s = 0.0;
    for (int i = 0; i < 14; ++i) s += g.weight[i] * es_f64(g.node[i]);
    return 8.0 * s;
}

float orc_es_f32(float x) { return es_f32(kBeta, x); }

void orc_cheb_seq_f32(float s, int n, float* out) { cheb_seq_f32(s, n, out); }
void orc_bessel_seq_f32(float alpha, int max_order, float* out) { bessel_seq_f32(alpha, max_order, out); }

void orc_jacobi3(const double* a, double* evals, double* evecs_colmajor) { jacobi3(a, evals, evecs_colmajor); }

// ------------------------------- preprocess ---------------------------------------------
// preprocess_measurement_set_gpu restated in FP32 (SVDI.cu:635-853): P1 expand, P2-P4
// mean/centre, P5 scatter, P6 eigen (FP64 Jacobi stands in for cusolverDnSsyevd), P7 basis,
// P8 project, P9 collapse.  Outputs Bin (S x 3 AoS), Vin (3x3 row-major), collapsed vis.
void orc_preprocess_ref32(size_t R, size_t C, const float* uvw, const float* freq,
                          const float* v0r, const float* v0i, const float* v3r, const float* v3i,
                          const uint8_t* f0, const uint8_t* f3, const float* w0, const float* w3,
                          float array_center_x, float* bin, float* vin, float* vis_re,
                          float* vis_im, float* evals_out) {
    const size_t S = R * C;
    std::vector<float> b(3 * S);
    for (size_t idx = 0; idx < S; ++idx) {                       // SVDI.cu:409-429
        const size_t row = idx % R, chan = idx / R;
        const float scale = freq[chan] / kSpeedOfLightF;
        for (int k = 0; k < 3; ++k) b[k * S + idx] = uvw[row * 3 + k] * scale;
    }
    float sums[3] = {0, 0, 0};                                   // SVDI.cu:431-465
    for (int k = 0; k < 3; ++k) {
        for (size_t blk = 0; blk < (S + 255) / 256; ++blk) {
            float loc[256];
            for (int t = 0; t < 256; ++t) {
                const size_t i = blk * 256 + t;
                loc[t] = i < S ? b[k * S + i] : 0.0f;
            }
            for (int stride = 128; stride > 0; stride >>= 1)
                for (int t = 0; t < stride; ++t) loc[t] += loc[t + stride];
            sums[k] += loc[0];
        }
        sums[k] /= static_cast<float>(S);                         // SVDI.cu:483-488
    }
    std::vector<float> cen(3 * S);                               // SVDI.cu:467-481
    for (int k = 0; k < 3; ++k)
        for (size_t i = 0; i < S; ++i) cen[k * S + i] = b[k * S + i] - sums[k];
    double cov[9];                                               // SVDI.cu:719-734 (Sgemm T,N)
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            float acc = 0.0f;
            for (size_t blk = 0; blk < S; blk += 4096) {
                float part = 0.0f;
                const size_t e = std::min(S, blk + 4096);
                for (size_t i = blk; i < e; ++i) part += cen[r * S + i] * cen[c * S + i];
                acc += part;
            }
            cov[r * 3 + c] = acc;
        }
    double ev[3], evec[9];
    jacobi3(cov, ev, evec);
    float evecf[9], basis[9];
    for (int i = 0; i < 9; ++i) evecf[i] = static_cast<float>(evec[i]);
    basis_from_evecs<float>(evecf, array_center_x, basis);
    for (int i = 0; i < 9; ++i) vin[i] = basis[i];
    for (int i = 0; i < 3; ++i) evals_out[i] = static_cast<float>(ev[i]);
    for (size_t idx = 0; idx < S; ++idx) {                       // SVDI.cu:509-536
        const float x0 = b[idx], x1 = b[S + idx], x2 = b[2 * S + idx];
        for (int r = 0; r < 3; ++r)
            bin[idx * 3 + r] = x0 * basis[r * 3 + 0] + x1 * basis[r * 3 + 1] + x2 * basis[r * 3 + 2];
    }
    for (size_t idx = 0; idx < S; ++idx) {                       // SVDI.cu:588-617
        const float a0 = f0[idx] ? 0.0f : (w0 ? w0[idx] : 1.0f);
        const float a3 = f3[idx] ? 0.0f : (w3 ? w3[idx] : 1.0f);
        const float ws = a0 + a3;
        if (ws > 0.0f) {
            vis_re[idx] = (v0r[idx] * a0 + v3r[idx] * a3) / ws;
            vis_im[idx] = (v0i[idx] * a0 + v3i[idx] * a3) / ws;
        } else {
            vis_re[idx] = 0.0f;
            vis_im[idx] = 0.0f;
        }
    }
}

// FP64 preprocess with the same semantics (mean-centred scatter, eigen, basis rules,
// uncentred projection), from FP64 uvw and frequencies.  Bin out is FP64 (S x 3).
void orc_preprocess_f64(size_t R, size_t C, const double* uvw, const double* freq,
                        double array_center_x, double* bin, double* vin, double* evals_out) {
    const size_t S = R * C;
    const double c = 299792458.0;
    long double sb[3] = {0, 0, 0};
    for (size_t idx = 0; idx < S; ++idx) {
        const size_t row = idx % R, chan = idx / R;
        const double sc = freq[chan] / c;
        for (int k = 0; k < 3; ++k) sb[k] += uvw[row * 3 + k] * sc;
    }
    double mean[3];
    for (int k = 0; k < 3; ++k) mean[k] = static_cast<double>(sb[k] / static_cast<long double>(S));
    long double m[9] = {0};
    for (size_t idx = 0; idx < S; ++idx) {
        const size_t row = idx % R, chan = idx / R;
        const double sc = freq[chan] / c;
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = uvw[row * 3 + k] * sc - mean[k];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) m[r * 3 + q] += d[r] * d[q];
    }
    double cov[9];
    for (int i = 0; i < 9; ++i) cov[i] = static_cast<double>(m[i]);
    double ev[3], evec[9];
    jacobi3(cov, ev, evec);
    basis_from_evecs<double>(evec, array_center_x, vin);
    for (int i = 0; i < 3; ++i) evals_out[i] = ev[i];
    for (size_t idx = 0; idx < S; ++idx) {
        const size_t row = idx % R, chan = idx / R;
        const double sc = freq[chan] / c;
        const double x0 = uvw[row * 3] * sc, x1 = uvw[row * 3 + 1] * sc, x2 = uvw[row * 3 + 2] * sc;
        for (int r = 0; r < 3; ++r)
            bin[idx * 3 + r] = x0 * vin[r * 3 + 0] + x1 * vin[r * 3 + 1] + x2 * vin[r * 3 + 2];
    }
}

// ------------------------------- planning -----------------------------------------------
// C1-C3: r3 range over samples and t3 stats over pixels (SVDI.cu:993-1110).
// out = {r3_min, r3_max, max_abs_t3, valid_pixels}.
void orc_cheb_stats(const float* bin, size_t S, const float* vin, size_t N, float cell, float* out) {
    float r3min = INFINITY, r3max = -INFINITY;
    for (size_t i = 0; i < S; ++i) {
        r3min = std::fmin(r3min, bin[i * 3 + 2]);
        r3max = std::fmax(r3max, bin[i * 3 + 2]);
    }
    float maxt3 = 0.0f;
    size_t valid = 0;
    const float v13 = vin[2], v23 = vin[5], v33 = vin[8];
    const float half = static_cast<float>(N) / 2.0f;
    for (size_t y = 0; y < N; ++y)
        for (size_t x = 0; x < N; ++x) {
            const float t1 = cell * (static_cast<float>(y) - half);
            const float t2 = cell * (static_cast<float>(x) - half);
            const float ra = 1.0f - (t1 + v13) * (t1 + v13) - (t2 + v23) * (t2 + v23);
            if (ra >= 0.0f) {
                const float bs = (v33 >= 0.0f) ? 1.0f : -1.0f;
                const float t3 = -v33 + bs * std::sqrt(ra);
                maxt3 = std::fmax(maxt3, std::fabs(t3));
                ++valid;
            }
        }
    out[0] = r3min;
    out[1] = r3max;
    out[2] = maxt3;
    out[3] = static_cast<float>(valid);
}

// finish_host_chebyshev_selection + make_chebyshev_slab_config (SVDI.cu:941-991).
// target_alpha <= 0 selects the default 10 (SVDI.cu:963-971).
// out = {alpha_max, target_alpha, num_slabs, r3_range, slab_width, slab_half, inv_half, slab_alpha}
void orc_slab_config(float r3min, float r3max, float max_abs_t3, float target_alpha, float* out) {
    const float half_range = std::fmax(0.5f * (r3max - r3min), 1.0e-6f);
    const float alpha_max = static_cast<float>(2.0f * kPi * max_abs_t3 * half_range);
    const float target = target_alpha > 0.0f ? target_alpha : 10.0f;
    const float range = std::fmax(r3max - r3min, 2.0e-6f);
    int ns = static_cast<int>(std::ceil(std::fmax(alpha_max, 0.0f) / target));
    if (ns < 1) ns = 1;
    const float width = range / static_cast<float>(ns);
    const float half = std::fmax(0.5f * width, 1.0e-6f);
    out[0] = alpha_max;
    out[1] = target;
    out[2] = static_cast<float>(ns);
    out[3] = range;
    out[4] = width;
    out[5] = half;
    out[6] = 1.0f / half;
    out[7] = static_cast<float>(2.0f * kPi * max_abs_t3 * half);
}

// C4 cheb_bucket_error_kernel restated in FP32 (SVDI.cu:1125-1202): max |e^{i a s} - sum| at
// K = 8, 16, 24, 32 over 257 alphas x 1024 s samples.
void orc_bucket_errors_ref32(float alpha_max, float* errs) {
    const int na = 257, ns = 1024;
    for (int b = 0; b < 4; ++b) errs[b] = 0.0f;
    for (int ai = 0; ai < na; ++ai) {
        const float alpha = alpha_max * static_cast<float>(ai) / static_cast<float>(na - 1);
        float j[kChebMax];
        bessel_seq_f32(alpha, kChebMax - 1, j);
        for (int si = 0; si < ns; ++si) {
            const float s = -1.0f + 2.0f * static_cast<float>(si) / static_cast<float>(ns - 1);
            const float er = std::cos(alpha * s), ei = std::sin(alpha * s);
            float ar = 0.0f, aiv = 0.0f, tm2 = 1.0f, tm1 = s;
            for (int n = 0; n < kChebMax; ++n) {
                float tn = 1.0f;
                if (n == 1) {
                    tn = tm1;
                } else if (n > 1) {
                    tn = 2.0f * s * tm1 - tm2;
                    tm2 = tm1;
                    tm1 = tn;
                }
                if (n == 0) {
                    ar += j[n] * tn;
                } else {
                    const float amp = 2.0f * j[n] * tn;
                    switch (n & 3) {
                        case 0: ar += amp; break;
                        case 1: aiv += amp; break;
                        case 2: ar -= amp; break;
                        default: aiv -= amp; break;
                    }
                }
                const int bucket = (n == 7) ? 0 : (n == 15) ? 1 : (n == 23) ? 2 : (n == 31) ? 3 : -1;
                if (bucket >= 0) {
                    const float dr = er - ar, di = ei - aiv;
                    errs[bucket] = std::fmax(errs[bucket], std::sqrt(dr * dr + di * di));
                }
            }
        }
    }
}

// Bucket choice of SVDI.cu:1577-1603: first of {8,16,24,32} with error <= 1e-2, else 32.
int orc_select_bucket(const float* errs) {
    const int terms[4] = {8, 16, 24, 32};
    for (int i = 0; i < 4; ++i)
        if (errs[i] <= kChebTargetError) return terms[i];
    return kChebMax;
}

// Archive K search, FIPk.cu:210-318: FP64 std::cyl_bessel_j, 513 alphas x 2048 s samples,
// smallest K in [1, 32] with sampled max error <= 1e-2.  errs_out[K-1] receives the sampled
// error of every K (the reference stops at the first passing K; the values it computes match).
int orc_archive_select_terms(float alpha_max, float* errs_out) {
    const int na = 513, ns = 2048;
    double maxerr[kChebMax] = {0};
    for (int ai = 0; ai < na; ++ai) {
        const float alpha_f = alpha_max * static_cast<float>(ai) / static_cast<float>(na - 1);
        const double alpha = static_cast<double>(alpha_f);
        double j[kChebMax];
        for (int n = 0; n < kChebMax; ++n) j[n] = std::cyl_bessel_j(static_cast<double>(n), alpha);
        for (int si = 0; si < ns; ++si) {
            const double s = -1.0 + 2.0 * static_cast<double>(si) / static_cast<double>(ns - 1);
            const std::complex<double> exact = std::exp(std::complex<double>(0.0, alpha * s));
            std::complex<double> sum(0.0, 0.0);
            double tm2 = 1.0, tm1 = s;
            for (int n = 0; n < kChebMax; ++n) {
                double tn;
                if (n == 0) tn = 1.0;
                else if (n == 1) tn = s;
                else {
                    tn = 2.0 * s * tm1 - tm2;
                    tm2 = tm1;
                    tm1 = tn;
                }
                if (n == 0) {
                    sum += std::complex<double>(j[n] * tn, 0.0);
                } else {
                    std::complex<double> ph;
                    switch (n & 3) {
                        case 0: ph = {1.0, 0.0}; break;
                        case 1: ph = {0.0, 1.0}; break;
                        case 2: ph = {-1.0, 0.0}; break;
                        default: ph = {0.0, -1.0}; break;
                    }
                    sum += 2.0 * ph * j[n] * tn;
                }
                maxerr[n] = std::max(maxerr[n], std::abs(exact - sum));
            }
        }
    }
    int chosen = -1;
    for (int k = 0; k < kChebMax; ++k) {
        if (errs_out) errs_out[k] = static_cast<float>(maxerr[k]);
        if (chosen < 0 && static_cast<float>(maxerr[k]) <= kChebTargetError) chosen = k + 1;
    }
    return chosen > 0 ? chosen : kChebMax;
}

// convolveKernel: SVDI.cu:1204-1216 (variant 0, cospif) and FIPk.cu:368-381 (variant 1,
// cosf(M_PI * ...)).  out has N/2+1 entries.
void orc_conv_corr(size_t N, size_t G, int variant, float* out) {
    const GL28& g = gl28();
    const float norm = 2.4937047051153827f;  // SVDI.cu:1512 (= 8 * sum w_i ES(x_i))
    for (size_t t = 0; t < N / 2 + 1; ++t) {
        const float tn = static_cast<float>(t) / G;
        float corr = 0.0f;
        for (int i = 0; i < 14; ++i) {
            const float node = static_cast<float>(g.node[i]);
            const float wgt = static_cast<float>(g.weight[i]);
            const float kern = static_cast<float>(es_f64(g.node[i]));
            float cs;
            if (variant == 0) {
                const float ang = tn * kSupport * node;
                cs = static_cast<float>(std::cos(kPi * static_cast<double>(ang)));
            } else {
                const float ang = static_cast<float>(kPi * tn * kSupport * node);
                cs = std::cos(ang);
            }
            corr += kern * wgt * cs;
        }
        out[t] = corr * kSupport / norm;
    }
}

// fused_gridding restated (SVDI.cu:1218-1293).  grid: planes x G x G complex interleaved,
// accumulated (caller zeroes).  Rows of the grid are split across threads; each thread scans
// every sample in order, so each cell sums its samples in sample order (deterministic).  A
// thread accumulates its rows in a band laid out cell-major ([row][column][plane], the planes of
// a cell adjacent) and copies it back at the end: the same additions in the same order as
// adding into the planes directly, without touching nslabs * K pages per cell.
void orc_grid_svdi(const float* bin, const float* vre, const float* vim, size_t S,
                   const float* vin, float r1r2_scale, size_t G, float r3_min, float slab_width,
                   float inv_half, int K, int slab_start, int nslabs, int total_slabs,
                   float* grid) {
    const long gmin = -static_cast<long>(G) / 2;
    const long gmax = (static_cast<long>(G) - 1) / 2;
    const long off = static_cast<long>(G) / 2;
    const float weight = std::fabs(vin[0] * vin[4] - vin[1] * vin[3]);
    const int nt = thread_count();
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
        const int nth = omp_get_num_threads();
#else
        const int tid = 0, nth = 1;
#endif
        const long row_lo = gmin + static_cast<long>(G) * tid / nth;
        const long row_hi = gmin + static_cast<long>(G) * (tid + 1) / nth - 1;
        const size_t P = static_cast<size_t>(nslabs) * K;
        const size_t band_rows = row_hi >= row_lo ? static_cast<size_t>(row_hi - row_lo + 1) : 0;
        std::vector<float> band(band_rows * G * P * 2);
        auto band_at = [&](long g1, long col, size_t plane) {   // (row g1, array column col)
            return band.data() + 2 * ((static_cast<size_t>(g1 - row_lo) * G + static_cast<size_t>(col)) * P + plane);
        };
        for (long g1 = row_lo; g1 <= row_hi; ++g1)
            for (size_t col = 0; col < G; ++col)
                for (size_t plane = 0; plane < P; ++plane) {
                    const float* src = grid + 2 * (plane * G * G + static_cast<size_t>(g1 + off) * G + col);
                    float* dst = band_at(g1, static_cast<long>(col), plane);
                    dst[0] = src[0];
                    dst[1] = src[1];
                }
        float cheb[kChebMax];
        float k1[16], k2[16];
        for (size_t idx = 0; idx < S; ++idx) {
            const float p1 = bin[idx * 3] * r1r2_scale;
            const float p2 = bin[idx * 3 + 1] * r1r2_scale;
            long a_lo = std::max(static_cast<long>(std::ceil(p1 - 4)), gmin);
            long a_hi = std::min(static_cast<long>(std::floor(p1 + 4)), gmax);
            if (a_hi < row_lo || a_lo > row_hi) continue;
            const float r3 = bin[idx * 3 + 2];
            int slab = static_cast<int>(std::floor((r3 - r3_min) / slab_width));
            slab = std::min(std::max(slab, 0), total_slabs - 1);
            if (slab < slab_start || slab >= slab_start + nslabs) continue;
            const float centre = r3_min + (static_cast<float>(slab) + 0.5f) * slab_width;
            float s = (r3 - centre) * inv_half;
            s = std::fmin(1.0f, std::fmax(-1.0f, s));
            cheb_seq_f32(s, K, cheb);
            const long b_lo = std::max(static_cast<long>(std::ceil(p2 - 4)), gmin);
            const long b_hi = std::min(static_cast<long>(std::floor(p2 + 4)), gmax);
            if (a_lo > a_hi || b_lo > b_hi) continue;
            for (long g = a_lo; g <= a_hi; ++g) k1[g - a_lo] = es_f32(kBeta, (g - p1) * 0.25f);
            for (long g = b_lo; g <= b_hi; ++g) k2[g - b_lo] = es_f32(kBeta, (g - p2) * 0.25f);
            const long ra = std::max(a_lo, row_lo), rb = std::min(a_hi, row_hi);
            for (long g1 = ra; g1 <= rb; ++g1)
                for (long g2 = b_lo; g2 <= b_hi; ++g2) {
                    float kv = k1[g1 - a_lo] * k2[g2 - b_lo];
                    if (((g1 + g2) & 1) != 0) kv = -kv;
                    for (int t = 0; t < K; ++t) {
                        const size_t plane = static_cast<size_t>(slab - slab_start) * K + t;
                        const float mw = (cheb[t] / weight) * kv;
                        float* c = band_at(g1, g2 + off, plane);
                        c[0] += vre[idx] * mw;
                        c[1] += vim[idx] * mw;
                    }
                }
        }
        for (long g1 = row_lo; g1 <= row_hi; ++g1)
            for (size_t col = 0; col < G; ++col)
                for (size_t plane = 0; plane < P; ++plane) {
                    const float* src = band_at(g1, static_cast<long>(col), plane);
                    float* dst = grid + 2 * (plane * G * G + static_cast<size_t>(g1 + off) * G + col);
                    dst[0] = src[0];
                    dst[1] = src[1];
                }
    }
}

// Archive gridding (FIPk.cu:383-452, with computeVisWeighted applied first): single slab,
// vis / |det| up front, weight = T_k * kernel.  Output complex interleaved like above.
void orc_grid_archive(const float* bin, const float* vre, const float* vim, size_t S,
                      float inten_scale, float uv_scale, size_t G, float r3_centre,
                      float inv_half, int K, float* grid) {
    const long gmin = -static_cast<long>(G) / 2;
    const long gmax = (static_cast<long>(G) - 1) / 2;
    const long off = static_cast<long>(G) / 2;
    const int nt = thread_count();
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
        const int nth = omp_get_num_threads();
#else
        const int tid = 0, nth = 1;
#endif
        const long row_lo = gmin + static_cast<long>(G) * tid / nth;
        const long row_hi = gmin + static_cast<long>(G) * (tid + 1) / nth - 1;
        float cheb[kChebMax];
        float k1[16], k2[16];
        for (size_t idx = 0; idx < S; ++idx) {
            const float p1 = bin[idx * 3] * uv_scale;
            const float p2 = bin[idx * 3 + 1] * uv_scale;
            const long a_lo = std::max(ceil_dev(p1 - 4), gmin);
            const long a_hi = std::min(floor_dev(p1 + 4), gmax);
            if (a_hi < row_lo || a_lo > row_hi) continue;
            const long b_lo = std::max(ceil_dev(p2 - 4), gmin);
            const long b_hi = std::min(floor_dev(p2 + 4), gmax);
            if (a_lo > a_hi || b_lo > b_hi) continue;
            const float vr = vre[idx] / inten_scale, vi = vim[idx] / inten_scale;
            float s = (bin[idx * 3 + 2] - r3_centre) * inv_half;
            s = std::fmin(1.0f, std::fmax(-1.0f, s));
            cheb_seq_f32(s, K, cheb);
            for (long g = a_lo; g <= a_hi; ++g) k1[g - a_lo] = es_f32(kBeta, (static_cast<float>(g) - p1) * 0.25f);
            for (long g = b_lo; g <= b_hi; ++g) k2[g - b_lo] = es_f32(kBeta, (static_cast<float>(g) - p2) * 0.25f);
            const long ra = std::max(a_lo, row_lo), rb = std::min(a_hi, row_hi);
            for (long g1 = ra; g1 <= rb; ++g1)
                for (long g2 = b_lo; g2 <= b_hi; ++g2) {
                    float kv = k1[g1 - a_lo] * k2[g2 - b_lo];
                    if (((g1 + g2) & 1) != 0) kv = -kv;
                    const long cell = (g1 + off) * static_cast<long>(G) + (g2 + off);
                    for (int t = 0; t < K; ++t) {
                        const float wgt = cheb[t] * kv;
                        float* c = grid + 2 * (static_cast<size_t>(t) * G * G + cell);
                        c[0] += vr * wgt;
                        c[1] += vi * wgt;
                    }
                }
        }
    }
}

// 2D inverse DFT (sign +, unnormalised) of P planes of G x G complex64, computed in FP64
// and rounded back to FP32 (stands in for cufftExecC2C(..., CUFFT_INVERSE), SVDI.cu:1713).
void orc_ifft2(float* data, size_t G, size_t P) {
    const int nt = thread_count();
    for (size_t p = 0; p < P; ++p) {
        float* plane = data + 2 * p * G * G;
        std::vector<std::complex<double>> buf(G * G);
        for (size_t i = 0; i < G * G; ++i) buf[i] = {plane[2 * i], plane[2 * i + 1]};
#pragma omp parallel num_threads(nt)
        {
            std::vector<std::complex<double>> line(G), work(G);
#pragma omp for
            for (long r = 0; r < static_cast<long>(G); ++r) {
                for (size_t c = 0; c < G; ++c) line[c] = buf[r * G + c];
                fft1d_inverse_f64(line.data(), G, work.data());
                for (size_t c = 0; c < G; ++c) buf[r * G + c] = line[c];
            }
#pragma omp for
            for (long c = 0; c < static_cast<long>(G); ++c) {
                for (size_t r = 0; r < G; ++r) line[r] = buf[r * G + c];
                fft1d_inverse_f64(line.data(), G, work.data());
                for (size_t r = 0; r < G; ++r) buf[r * G + c] = line[r];
            }
        }
        for (size_t i = 0; i < G * G; ++i) {
            plane[2 * i] = static_cast<float>(buf[i].real());
            plane[2 * i + 1] = static_cast<float>(buf[i].imag());
        }
    }
}

// 1D inverse DFT in FP64 (for pinning the FFT against numpy).
void orc_ifft1_f64(double* data_interleaved, size_t n) {
    std::vector<std::complex<double>> buf(n), work(n);
    for (size_t i = 0; i < n; ++i) buf[i] = {data_interleaved[2 * i], data_interleaved[2 * i + 1]};
    fft1d_inverse_f64(buf.data(), n, work.data());
    for (size_t i = 0; i < n; ++i) {
        data_interleaved[2 * i] = buf[i].real();
        data_interleaved[2 * i + 1] = buf[i].imag();
    }
}

// Reprojection maps for every pixel: out[(y*N + x)*2 + {0,1}] = (oi0, oi1).
void orc_reproject(size_t N, float cell, const float* vin, int variant, float* out) {
    for (size_t y = 0; y < N; ++y)
        for (size_t x = 0; x < N; ++x) {
            float a, b;
            if (variant == 0) reproject_svdi(N, cell, vin, x, y, &a, &b);
            else reproject_archive(N, cell, vin, x, y, &a, &b);
            out[(y * N + x) * 2] = a;
            out[(y * N + x) * 2 + 1] = b;
        }
}

// Bilinear splat of a full image of values through a reprojection map (SVDI.cu:1447-1461).
void orc_splat(size_t N, const float* map, const float* values, float* dirty) {
    for (size_t p = 0; p < N * N; ++p) splat(dirty, N, map[2 * p], map[2 * p + 1], values[p]);
}

// fused_interpolation restated (SVDI.cu:1295-1463) for one slab batch.  planes: batch planes
// of G x G complex after the inverse FFT.  dirty accumulates the splat (caller zeroes).
// signed_out (nullable) receives the signed, grid-corrected pixel sum times 1/S at (x, y).
void orc_image_svdi(const float* planes, size_t G, size_t N, float cell, const float* vin,
                    const float* conv, float norm, float invS, float r3_min, float slab_width,
                    float slab_half, int K, int slab_start, int nslabs, float* dirty,
                    float* signed_out) {
    const long half = static_cast<long>(N / 2);
    std::vector<float> vals(N * N, 0.0f), map(2 * N * N);
    std::vector<uint8_t> valid(N * N, 0);
    const int nt = thread_count();
#pragma omp parallel for num_threads(nt)
    for (long y = 0; y < static_cast<long>(N); ++y) {
        float j[kChebMax];
        for (size_t x = 0; x < N; ++x) {
            float oi0, oi1;
            reproject_svdi(N, cell, vin, x, y, &oi0, &oi1);
            map[2 * (y * N + x)] = oi0;
            map[2 * (y * N + x) + 1] = oi1;
            const long idxr = static_cast<long>(x) - half;
            const long idyr = static_cast<long>(y) - half;
            const float t1 = cell * static_cast<float>(idyr);
            const float t2 = cell * static_cast<float>(idxr);
            const float v13 = vin[2], v23 = vin[5], v33 = vin[8];
            const float ra = 1.0f - (t1 + v13) * (t1 + v13) - (t2 + v23) * (t2 + v23);
            if (ra < 0.0f) continue;
            const float bs = (v33 >= 0.0f) ? 1.0f : -1.0f;
            const float t3 = -v33 + bs * std::sqrt(ra);
            const float alpha = static_cast<float>(2.0f * kPi * t3 * slab_half);
            bessel_seq_f32(alpha, K - 1, j);
            float sum = 0.0f;
            const long gidx = static_cast<long>(G * G / 2 + G / 2) + idyr * static_cast<long>(G) + idxr;
            for (int ls = 0; ls < nslabs; ++ls) {
                const int slab = slab_start + ls;
                const float centre = r3_min + (static_cast<float>(slab) + 0.5f) * slab_width;
                const float beta = static_cast<float>(2.0f * kPi * t3 * centre);
                const float cb = std::cos(beta), sb = std::sin(beta);
                for (int t = 0; t < K; ++t) {
                    const size_t plane = static_cast<size_t>(ls) * K + t;
                    const float* m = planes + 2 * (plane * G * G + gidx);
                    float cr, ci;
                    ja_coeff(t, j[t], cb, sb, &cr, &ci);
                    sum += cr * m[0] - ci * m[1];
                }
            }
            if ((idxr + idyr) & 1) sum = -sum;
            sum *= 1 / (conv[std::labs(idxr)] * conv[std::labs(idyr)] * norm * norm);
            if (signed_out) signed_out[y * N + x] = sum * invS;
            vals[y * N + x] = std::fabs(sum) * invS;
            valid[y * N + x] = 1;
        }
    }
    for (size_t p = 0; p < N * N; ++p)
        if (valid[p]) splat(dirty, N, map[2 * p], map[2 * p + 1], vals[p]);
}

// Archive image pass: accumulation + scaling + coordschange/p2p + finalinterp
// (FIPk.cu:463-677).  Single slab centred at r3_centre with half range r3_half.
void orc_image_archive(const float* planes, size_t G, size_t N, float cell, const float* vin,
                       const float* conv, float norm, size_t num_baselines, float r3_centre,
                       float r3_half, int K, float* dirty, float* signed_out) {
    const long half = static_cast<long>(N / 2);
    std::vector<float> pre(N * N, 0.0f), map(2 * N * N);
    const int nt = thread_count();
#pragma omp parallel for num_threads(nt)
    for (long yy = 0; yy < static_cast<long>(N); ++yy) {
        float j[kChebMax];
        for (long xx = 0; xx < static_cast<long>(N); ++xx) {
            float a, b;
            reproject_archive(N, cell, vin, xx, yy, &a, &b);
            map[2 * (yy * N + xx)] = a;
            map[2 * (yy * N + xx) + 1] = b;
            const long idx = xx - half, idy = yy - half;
            const float t1 = cell * static_cast<float>(idy);
            const float t2 = cell * static_cast<float>(idx);
            const float v13 = vin[2], v23 = vin[5], v33 = vin[8];
            const float ra = 1.0f - (t1 + v13) * (t1 + v13) - (t2 + v23) * (t2 + v23);
            if (ra < 0.0f) continue;
            const float bs = (v33 >= 0.0f) ? 1.0f : -1.0f;
            const float t3 = -v33 + bs * std::sqrt(ra);
            const float alpha = static_cast<float>(2.0f * kPi * t3 * r3_half);
            const float beta = static_cast<float>(2.0f * kPi * t3 * r3_centre);
            const float cb = std::cos(beta), sb = std::sin(beta);
            bessel_seq_f32(alpha, K - 1, j);
            float sum = 0.0f;
            const size_t gofs = G * G / 2 + G / 2;
            for (int t = 0; t < K; ++t) {
                const float* m = planes + 2 * (static_cast<size_t>(t) * G * G + gofs + idy * static_cast<long>(G) + idx);
                float cr, ci;
                ja_coeff(t, j[t], cb, sb, &cr, &ci);
                sum += cr * m[0] - ci * m[1];
            }
            if (((std::labs(idx) + std::labs(idy)) & 1) != 0) sum = -sum;
            pre[yy * N + xx] = sum;
        }
    }
    const float inv_nb = 1.0f / static_cast<float>(num_baselines);
    for (long yy = 0; yy < static_cast<long>(N); ++yy)
        for (long xx = 0; xx < static_cast<long>(N); ++xx) {
            const long idx = xx - half, idy = yy - half;
            float v = pre[yy * N + xx];
            v *= 1 / (conv[std::labs(idx)] * conv[std::labs(idy)] * norm * norm);
            if (signed_out) signed_out[yy * N + xx] = v * inv_nb;
            pre[yy * N + xx] = std::fabs(v);
        }
    for (size_t p = 0; p < N * N; ++p) splat(dirty, N, map[2 * p], map[2 * p + 1], pre[p] * inv_nb);
}

// ------------------------------- FP64 DFT -----------------------------------------------
// I(x, y) = Re sum_j keep_j V_j exp(+2 pi i (r1 t1 + r2 t2 + r3 t3)) / (|V00 V11 - V01 V10| S),
// t1 = cell (y - N/2), t2 = cell (x - N/2), t3 = -v33 + sgn(v33) sqrt(1 - (t1+v13)^2 - (t2+v23)^2).
// pixels: npix (x, y) pairs.  Pixels with a negative root argument return NaN.
void orc_dft64(const double* bin, const double* vre, const double* vim, const uint8_t* keep,
               size_t S, const double* vin, size_t N, double cell, const int32_t* pixels,
               size_t npix, double* out) {
    const double det = std::fabs(vin[0] * vin[4] - vin[1] * vin[3]);
    const long half = static_cast<long>(N / 2);
    const int nt = thread_count();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (long p = 0; p < static_cast<long>(npix); ++p) {
        const long x = pixels[2 * p], y = pixels[2 * p + 1];
        const double t1 = cell * static_cast<double>(y - half);
        const double t2 = cell * static_cast<double>(x - half);
        const double ra = 1.0 - (t1 + vin[2]) * (t1 + vin[2]) - (t2 + vin[5]) * (t2 + vin[5]);
        if (ra < 0.0) {
            out[p] = NAN;
            continue;
        }
        const double bs = vin[8] >= 0.0 ? 1.0 : -1.0;
        const double t3 = -vin[8] + bs * std::sqrt(ra);
        double acc = 0.0, comp = 0.0;
        for (size_t j = 0; j < S; ++j) {
            if (keep && !keep[j]) continue;
            const double ph = 2.0 * kPi * (bin[3 * j] * t1 + bin[3 * j + 1] * t2 + bin[3 * j + 2] * t3);
            const double term = vre[j] * std::cos(ph) - vim[j] * std::sin(ph);
            const double yk = term - comp;             // Kahan summation
            const double tk = acc + yk;
            comp = (tk - acc) - yk;
            acc = tk;
        }
        out[p] = acc / (det * static_cast<double>(S));
    }
}

int orc_num_threads() { return thread_count(); }

}  // extern "C"
