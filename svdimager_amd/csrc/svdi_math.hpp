// Scalar math shared by the HIP kernels and by CPU probes (the same source compiles for both).
//
// Exponential-of-semicircle (ES) kernel, Chebyshev recurrence, Bessel J_n(x) in FP64, the 3x3
// symmetric eigen-solver that replaces cusolverDnSsyevd, the reference's basis orientation
// rules, and the SIN reprojection of one pixel.  Each function cites the reference lines whose
// behaviour it reproduces.
#pragma once

#include <cmath>
#include <cstdint>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

#define SVDI_HD __host__ __device__ __forceinline__

namespace svdi {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSpeedOfLight = 299792458.0;
constexpr float kBeta = 15.3704324328f;          // ES shape, SVDIntegratedImager.cu:1235
constexpr int kSupport = 8;                      // ES support, SVDIntegratedImager.cu:1234
constexpr int kMaxTerms = 32;                    // CHEB_MAX_MOMENTS, SVDIntegratedImager.cu:35
constexpr float kTargetError = 1.0e-2f;          // CHEB_TARGET_ERROR, SVDIntegratedImager.cu:36
constexpr float kDefaultTargetAlpha = 10.0f;     // CHEB_TARGET_ALPHA, SVDIntegratedImager.cu:37
constexpr float kConvCorrNorm = 2.4937047051153827f;  // 8 * sum_i w_i ES(x_i), SVDIntegratedImager.cu:1512

// exp(beta (sqrt(1 - x^2) - 1)) for |x| <= 1, else 0 (SVDIntegratedImager.cu:884-888).
SVDI_HD float es_kernel(float x) {
    const float xx = x * x;
    return (xx > 1.0f) ? 0.0f : expf(kBeta * (sqrtf(1.0f - xx) - 1.0f));
}

// The same kernel with the hardware square root and base-2 exponential on the device (about
// 1 ulp each, no denormal fix-ups: ES >= e^-beta > 1e-7 wherever it is non-zero).
SVDI_HD float es_kernel_fast(float x) {
    const float xx = x * x;
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_sqrtf(1.0f - xx);
    const float e = __builtin_amdgcn_exp2f((kBeta * 1.4426950408889634f) * (r - 1.0f));
#else
    const float r = sqrtf(1.0f - xx);
    const float e = exp2f((kBeta * 1.4426950408889634f) * (r - 1.0f));
#endif
    return (xx > 1.0f) ? 0.0f : e;
}

SVDI_HD double es_kernel_f64(double x) {
    const double xx = x * x;
    return (xx > 1.0) ? 0.0 : exp(15.3704324328 * (sqrt(1.0 - xx) - 1.0));
}

// J_n(x) by its power series; accurate in FP64 for |x| <= 20.
SVDI_HD double bessel_series(int n, double x) {
    const double h = 0.5 * x;
    double term = 1.0;
    for (int m = 1; m <= n; ++m) term *= h / m;
    double sum = term;
    const double q = -h * h;
    for (int k = 1; k < 200; ++k) {
        term *= q / (static_cast<double>(k) * static_cast<double>(n + k));
        sum += term;
        if (fabs(term) <= 1e-17 * fabs(sum)) break;
    }
    return sum;
}

// J_0(x), J_1(x) for x >= 1000 from Hankel's asymptotic expansion (Abramowitz & Stegun
// 9.2.5-9.2.10), truncated after the x^-3 terms: the next terms are below 1e-13 there.
SVDI_HD void bessel_j01_asymptotic(double x, double* j0, double* j1) {
    const double ix = 1.0 / x, ix2 = ix * ix;
    const double p0 = 1.0 - 9.0 / 128.0 * ix2;
    const double q0 = -0.125 * ix + 75.0 / 1024.0 * ix2 * ix;
    const double p1 = 1.0 + 15.0 / 128.0 * ix2;
    const double q1 = 0.375 * ix - 105.0 / 1024.0 * ix2 * ix;
    const double amp = sqrt(2.0 / (kPi * x));
    const double c0 = x - 0.25 * kPi, c1 = x - 0.75 * kPi;
    *j0 = amp * (p0 * cos(c0) - q0 * sin(c0));
    *j1 = amp * (p1 * cos(c1) - q1 * sin(c1));
}

// Below this |x| the power series converges in a few terms (|x/2|^2 <= 2.5e-5 per term); above
// it Miller's recurrence is cheaper and as accurate.
constexpr double kBesselSeriesMax = 1e-2;

// J_0 .. J_{count-1}(x) in FP64.  |x| < 1e-2: the power series per order; |x| >= 1000: forward
// recurrence from the asymptotic J_0, J_1 (stable while n < x); otherwise Miller's backward
// recurrence normalised by J_0 + 2 sum_k J_{2k} = 1.  Replaces the FP32 power series of
// bessel_sequence (SVDIntegratedImager.cu:901-920), which loses digits to cancellation near x = 10.
// J_0..J_{COUNT-1}(x) with every index known at compile time, so the sequence stays in
// registers on the device: Miller's backward recurrence runs from m down to COUNT + 1 without
// stores, then through the COUNT stored terms unrolled.  As the runtime-count form below, except
// that for COUNT <= 16 the recurrence starts at a fixed m and runs fully unrolled: for
// 1e-2 <= |x| < 2 at m = max(COUNT, 8) + 12 (even), where J_m(x) / J_{COUNT-1}(x) < 1e-15 and the
// recurrence grows by less than (2m / |x|)^m < 1e110; for 2 <= |x| < 8 at m = 36, where
// J_37(x) < 1e-21, J_37(x) / J_15(x) < 4e-18 and the growth is at most J_0(x) / J_36(x) < 1e41.
// Neither needs rescaling.
constexpr double kBesselFixedMax = 8.0;
constexpr int kBesselFixedStart = 36;

// Miller's backward recurrence from J_M = 1, J_{M+1} = 0 (M even), unrolled, normalised by
// J_0 + 2 sum_k J_{2k} = 1: J_0..J_{COUNT-1}(ax), ax > 0.
template <int COUNT, int M>
SVDI_HD void bessel_miller_fixed(double ax, double* j) {
    static_assert(M % 2 == 0 && M > COUNT, "start order");
    const double two_over_x = 2.0 / ax;
    double jp = 0.0, jc = 1.0, norm = 0.0;
#pragma unroll
    for (int k = M; k > COUNT; --k) {
        const double jm = (k * two_over_x) * jc - jp;   // J_{k-1}
        jp = jc;
        jc = jm;
        if (((k - 1) & 1) == 0) norm += 2.0 * jc;
    }
#pragma unroll
    for (int k = COUNT; k > 0; --k) {
        const double jm = (k * two_over_x) * jc - jp;
        jp = jc;
        jc = jm;
        j[k - 1] = jc;
        if (k - 1 > 0 && ((k - 1) & 1) == 0) norm += 2.0 * jc;
    }
    norm += jc;                                        // + J_0
    const double inv = 1.0 / norm;
#pragma unroll
    for (int n = 0; n < COUNT; ++n) j[n] *= inv;
}

template <int COUNT>
SVDI_HD void bessel_sequence_fixed(double x, double* j) {
    const double ax = fabs(x);
    if (ax < kBesselSeriesMax) {
        // J_n = h^n / n! (1 + q r1 (1 + q r2 (1 + q r3 (1 + q r4)))), h = x / 2, q = -h^2,
        // r_k = 1 / (k (n + k)): the power series through q^4 (|q| < 2.5e-5, so the q^5 term is
        // below 1e-25 of the first), with the reciprocals folded at compile time.
        const double h = 0.5 * ax, q = -h * h;
        double p = 1.0;                                    // h^n / n!
#pragma unroll
        for (int n = 0; n < COUNT; ++n) {
            const double r1 = 1.0 / (1.0 * (n + 1)), r2 = 1.0 / (2.0 * (n + 2));
            const double r3 = 1.0 / (3.0 * (n + 3)), r4 = 1.0 / (4.0 * (n + 4));
            const double s = fma(q * r1, fma(q * r2, fma(q * r3, fma(q, r4, 1.0), 1.0), 1.0), 1.0);
            j[n] = p * s;
            p *= h * (1.0 / (n + 1));
        }
    } else if (COUNT <= 16 && ax < 2.0) {
        bessel_miller_fixed<COUNT, 2 * (((COUNT > 8 ? COUNT : 8) + 13) / 2)>(ax, j);
    } else if (COUNT <= 16 && ax < kBesselFixedMax) {
        bessel_miller_fixed<COUNT, kBesselFixedStart>(ax, j);
    } else if (ax >= 1000.0) {
        double j0, j1;
        bessel_j01_asymptotic(ax, &j0, &j1);
        j[0] = j0;
        if (COUNT > 1) j[1] = j1;
#pragma unroll
        for (int n = 1; n + 1 < COUNT; ++n) j[n + 1] = (2.0 * n / ax) * j[n] - j[n - 1];
    } else {
        const double big = 1e150, small = 1e-150;
        const int top = static_cast<int>(fmax(static_cast<double>(COUNT), ax));
        const int m = 2 * ((top + 20 + static_cast<int>(sqrt(60.0 * top))) / 2);   // > COUNT
        const double two_over_x = 2.0 / ax;
        double jp = 0.0, jc = 1.0, norm = 0.0;
#pragma unroll
        for (int n = 0; n < COUNT; ++n) j[n] = 0.0;
        for (int k = m; k > COUNT; --k) {
            const double jm = (k * two_over_x) * jc - jp;   // J_{k-1}
            jp = jc;
            jc = jm;
            if (fabs(jc) > big) {
                jc *= small;
                jp *= small;
                norm *= small;
            }
            const int km1 = k - 1;
            if ((km1 & 1) == 0) norm += 2.0 * jc;
        }
#pragma unroll
        for (int k = COUNT; k > 0; --k) {
            const double jm = (k * two_over_x) * jc - jp;
            jp = jc;
            jc = jm;
            if (fabs(jc) > big) {
                jc *= small;
                jp *= small;
                norm *= small;
#pragma unroll
                for (int n = 0; n < COUNT; ++n) j[n] *= small;
            }
            j[k - 1] = jc;
            if (k - 1 > 0 && ((k - 1) & 1) == 0) norm += 2.0 * jc;
        }
        norm += jc;                                        // + J_0
        const double inv = 1.0 / norm;
#pragma unroll
        for (int n = 0; n < COUNT; ++n) j[n] *= inv;
    }
    if (x < 0.0) {
#pragma unroll
        for (int n = 1; n < COUNT; n += 2) j[n] = -j[n];
    }
}

SVDI_HD void bessel_sequence_f64(double x, int count, double* j) {
    const double ax = fabs(x);
    if (ax < kBesselSeriesMax) {
        for (int n = 0; n < count; ++n) j[n] = bessel_series(n, ax);
    } else if (ax >= 1000.0) {
        double j0, j1;
        bessel_j01_asymptotic(ax, &j0, &j1);
        if (count > 0) j[0] = j0;
        if (count > 1) j[1] = j1;
        for (int n = 1; n + 1 < count; ++n) j[n + 1] = (2.0 * n / ax) * j[n] - j[n - 1];
    } else {
        const double big = 1e150, small = 1e-150;
        const int top = static_cast<int>(fmax(static_cast<double>(count), ax));
        int m = 2 * ((top + 20 + static_cast<int>(sqrt(60.0 * top))) / 2);
        const double two_over_x = 2.0 / ax;
        double jp = 0.0, jc = 1.0, norm = 0.0;
        for (int n = 0; n < count; ++n) j[n] = 0.0;
        for (int k = m; k > 0; --k) {
            const double jm = (k * two_over_x) * jc - jp;   // J_{k-1}
            jp = jc;
            jc = jm;
            if (fabs(jc) > big) {
                jc *= small;
                jp *= small;
                norm *= small;
                for (int n = 0; n < count; ++n) j[n] *= small;
            }
            const int km1 = k - 1;
            if (km1 < count) j[km1] = jc;
            if (km1 > 0 && (km1 & 1) == 0) norm += 2.0 * jc;
        }
        norm += jc;                                        // + J_0
        const double inv = 1.0 / norm;
        for (int n = 0; n < count; ++n) j[n] *= inv;
    }
    if (x < 0.0)
        for (int n = 1; n < count; n += 2) j[n] = -j[n];
}

// Chebyshev T_0..T_{count-1}(s) by the three-term recurrence (SVDIntegratedImager.cu:890-899).
template <typename T>
SVDI_HD void chebyshev_sequence(T s, int count, T* t) {
    t[0] = T(1);
    if (count > 1) t[1] = s;
    for (int k = 2; k < count; ++k) t[k] = T(2) * s * t[k - 1] - t[k - 2];
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations in FP64.  Eigenvalues
// ascending; eigenvectors returned as columns of a column-major matrix, the convention of
// cusolverDnSsyevd (SVDIntegratedImager.cu:736-760).
SVDI_HD void eigen_sym3(const double* a_in, double* evals, double* evecs_colmajor) {
    double a[3][3], v[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            a[r][c] = 0.5 * (a_in[r * 3 + c] + a_in[c * 3 + r]);
            v[r][c] = (r == c) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        const double diag = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off <= 1e-300 || off <= 1e-17 * diag) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = a[p][q];
                // An off-diagonal element below the rounding level of its diagonal pair is set
                // to zero, so the sweep loop ends once the matrix is diagonal to working
                // precision (relative eigenvalue change <= 1e-17).
                if (fabs(apq) <= 1e-17 * (fabs(a[p][p]) + fabs(a[q][q]))) {
                    a[p][q] = 0.0;
                    a[q][p] = 0.0;
                    continue;
                }
                const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0);
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
    int order[3] = {0, 1, 2};
    for (int i = 0; i < 2; ++i)                          // sort ascending by eigenvalue
        for (int k = 0; k < 2 - i; ++k)
            if (a[order[k]][order[k]] > a[order[k + 1]][order[k + 1]]) {
                const int tmp = order[k];
                order[k] = order[k + 1];
                order[k + 1] = tmp;
            }
    for (int j = 0; j < 3; ++j) {
        evals[j] = a[order[j]][order[j]];
        for (int r = 0; r < 3; ++r) evecs_colmajor[j * 3 + r] = v[r][order[j]];
    }
}

// Basis rows = eigenvectors by descending eigenvalue (build_basis_kernel,
// SVDIntegratedImager.cu:490-507), oriented so R00 > 0, R11 * array_center_x >= 0 when
// array_center_x != 0, and R2 = R0 x R1 normalised (SVDIntegratedImager.cu:538-586).
SVDI_HD void oriented_basis(const double* evecs_colmajor, double array_center_x, double* basis) {
    for (int axis = 0; axis < 3; ++axis) {
        basis[0 * 3 + axis] = evecs_colmajor[axis + 2 * 3];
        basis[1 * 3 + axis] = evecs_colmajor[axis + 1 * 3];
        basis[2 * 3 + axis] = evecs_colmajor[axis + 0 * 3];
    }
    if (basis[0] < 0.0)
        for (int a = 0; a < 3; ++a) basis[a] = -basis[a];
    if (array_center_x != 0.0 && basis[4] * array_center_x < 0.0)
        for (int a = 0; a < 3; ++a) basis[3 + a] = -basis[3 + a];
    const double r00 = basis[0], r01 = basis[1], r02 = basis[2];
    const double r10 = basis[3], r11 = basis[4], r12 = basis[5];
    double r20 = r01 * r12 - r02 * r11;
    double r21 = r02 * r10 - r00 * r12;
    double r22 = r00 * r11 - r01 * r10;
    const double norm = sqrt(r20 * r20 + r21 * r21 + r22 * r22);
    if (norm > 0.0) {
        r20 /= norm;
        r21 /= norm;
        r22 /= norm;
    }
    basis[6] = r20;
    basis[7] = r21;
    basis[8] = r22;
}

// Target of the bilinear splat for pixel (idx, idy): the rotated-frame pixel is mapped to the
// sky frame and re-projected with the SIN (slant orthographic) projection of Calabretta &
// Greisen (2002) with obliquity (xi, eta) = (V20, V21) / V22.  FP32 operation order of
// SVDIntegratedImager.cu:1338-1379.
SVDI_HD void sin_reproject(long di, float dc_rad, const float* V, long idx, long idy,
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
    const float p1 = (-V00 * idxf + V10 * idyf) / fabsf(V22) + di2;
    const float p2 = (-V01 * idxf + V11 * idyf) / fabsf(V22) + di2;
    float x = -dc * (p1 - (di2 + 1.0f));
    float y = dc * (p2 - (di2 + 1.0f));
    const float h = hypotf(x, y);
    const float hr0 = h / r0;
    if (h != 0.0f) {
        x /= h;
        y /= -h;
    } else {
        x = 0.0f;
        y = 1.0f;
    }
    if (h <= r0) {
        const float r = h;
        float z;
        if (h < r0 * sqrtf(0.5f)) {
            z = 1.0f - sqrtf(1.0f - hr0 * hr0);
        } else {
            z = hr0 * hr0 / (1.0f + sqrtf(1.0f - hr0 * hr0));
        }
        const float w = xi * xi + eta * eta;
        if (w == 0.0f) {
            x = r * x;
            y = -r * y;
        } else {
            x = r * x + z * r0 * xi;
            y = -r * y + z * r0 * eta;
        }
        *oi0_out = -x / dc + di2 + 1.0f;
        *oi1_out = y / dc + di2 + 1.0f;
    } else {
        *oi0_out = p1;
        *oi1_out = p2;
    }
}

// Degree-based trig helpers of the archive imager (archive/FIPkernels.cu:50-166), FP32: exact
// values at multiples of 90 degrees, otherwise the float functions of the angle in radians.
SVDI_HD long floor_toward(float v) {
    return v < 0 ? -static_cast<long>(ceilf(-v)) : static_cast<long>(floorf(v));
}
SVDI_HD void sincosd(float angle, float& s, float& c) {
    if (fmodf(angle, 90.0f) == 0) {
        const int i = static_cast<int>(fabsf(static_cast<float>(floor_toward(angle / 90.0f + 0.5f)))) % 4;
        switch (i) {
            case 0: s = 0.0f; c = 1.0f; break;
            case 1: s = (angle > 0.0f) ? 1.0f : -1.0f; c = 0.0f; break;
            case 2: s = 0.0f; c = -1.0f; break;
            default: s = (angle > 0.0f) ? -1.0f : 1.0f; c = 0.0f; break;
        }
    } else {
        s = sinf(static_cast<float>(angle * kPi / 180.0f));
        c = cosf(static_cast<float>(angle * kPi / 180.0f));
    }
}
SVDI_HD float sind(float angle) {
    if (fmodf(angle, 90.0f) == 0) {
        const int i = static_cast<int>(fabsf(static_cast<float>(floor_toward(angle / 90.0f - 0.5f)))) % 4;
        return i == 0 ? 1.0f : (i == 2 ? -1.0f : 0.0f);
    }
    return sinf(static_cast<float>(angle * kPi / 180.0f));
}
SVDI_HD float cosd(float angle) {
    if (fmodf(angle, 90.0f) == 0) {
        const int i = static_cast<int>(fabsf(static_cast<float>(floor_toward(angle / 90.0f + 0.5f)))) % 4;
        return i == 0 ? 1.0f : (i == 2 ? -1.0f : 0.0f);
    }
    return cosf(static_cast<float>(angle * kPi / 180.0f));
}
SVDI_HD float atan2d(float y, float x) {
    if (y == 0.0f) return (x >= 0.0f) ? 0.0f : 180.0f;
    if (x == 0.0f) return (y > 0.0f) ? 90.0f : -90.0f;
    return static_cast<float>(atan2f(y, x) * 180.0f / kPi);
}
SVDI_HD float acosd(float v) {
    if (v >= 1.0f && v - 1.0f < 1e-10) return 0.0f;
    if (v == 0.0f) return 90.0f;
    if (v <= -1.0f && v + 1.0f > -1e-10) return 180.0f;
    return static_cast<float>(acosf(v) * 180.0f / kPi);
}
SVDI_HD float asind(float v) {
    if (v <= -1.0f && v + 1.0f > -1e-10) return -90.0f;
    if (v == 0.0f) return 0.0f;
    if (v >= 1.0f && v - 1.0f < 1e-10) return 90.0f;
    return static_cast<float>(asinf(v) * 180.0f / kPi);
}

// The archive's two-step form of the same re-projection: coordschange then p2p
// (archive/FIPkernels.cu:537-640), FP32 with degree-based trig.
SVDI_HD void sin_reproject_archive(long di, float dc_in, const float* V, long idx, long idy,
                                   float* out0, float* out1) {
    const float half = static_cast<float>(di / 2);
    const float p1 = (-V[0] * (static_cast<float>(idx) - half) + V[3] * (static_cast<float>(idy) - half)) /
                         fabsf(V[8]) + half;
    const float p2 = (-V[1] * (static_cast<float>(idx) - half) + V[4] * (static_cast<float>(idy) - half)) /
                         fabsf(V[8]) + half;
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
        theta = acosd(sqrtf(r2));
    } else if (r2 <= 1.0f) {
        theta = asind(sqrtf(1.0f - r2));
    } else {
        return;
    }
    float sinphi, cosphi;
    sincosd(phi, sinphi, cosphi);
    x = sinphi;
    y = cosphi;
    const float t = static_cast<float>((90.0f - fabsf(theta)) / 180.0f * kPi);
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

}  // namespace svdi
