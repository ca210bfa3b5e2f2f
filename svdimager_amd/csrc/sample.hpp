// Per-sample model shared by the gridders (fused_gridding, SVDIntegratedImager.cu:1218-1293;
// gridding, archive/FIPkernels.cu:391-452): position, clipped footprint, slab, Chebyshev
// argument and weighted visibility of one sample, from either input form.
#pragma once

#include "svdi_internal.hpp"

namespace svdi {

struct Footprint {
    int a0, na, b0, nb;     // first grid row / column (array coordinates) and extents (<= 9)
    float da, db;           // a0 - p1 and b0 - p2 (centred coordinates), cells
    float q1, q2;           // preprocessed input: p1, p2 (FP32); snapshot input: p - a0 = -da
    int local_slab;
    float s;                // Chebyshev argument in [-1, 1]
    float2 vw;              // visibility / |det|
};

template <bool kSnap>
__device__ __forceinline__ bool setup_sample(const SampleInput& in, const GridParams& gp,
                                             size_t idx, size_t row, size_t chan, Footprint& f) {
    const long gmin = -static_cast<long>(gp.G) / 2;
    const long gmax = (static_cast<long>(gp.G) - 1) / 2;
    const long off = static_cast<long>(gp.G) / 2;
    float2 v;
    long lo1, hi1, lo2, hi2;
    float da, db, s;
    int slab = 0;
    if constexpr (kSnap) {
        const double sc = in.freq[chan] / kSpeedOfLight;
        const double p1 = in.rot[3 * row] * sc * gp.scale_f64;
        const double p2 = in.rot[3 * row + 1] * sc * gp.scale_f64;
        const double r3 = in.rot[3 * row + 2] * sc;
        double centre;
        if (gp.archive) {
            centre = gp.r3_centre;
        } else {
            slab = static_cast<int>(floor((r3 - static_cast<double>(gp.r3_min)) / gp.slab_width));
            slab = min(max(slab, 0), gp.total_slabs - 1);
            if (slab < gp.slab_start || slab >= gp.slab_start + gp.nslabs) return false;
            centre = gp.r3_min + (static_cast<float>(slab) + 0.5f) * gp.slab_width;
        }
        s = static_cast<float>(fmin(1.0, fmax(-1.0, (r3 - centre) * static_cast<double>(gp.inv_half))));
        lo1 = max(static_cast<long>(ceil(p1 - 4.0)), gmin);
        hi1 = min(static_cast<long>(floor(p1 + 4.0)), gmax);
        lo2 = max(static_cast<long>(ceil(p2 - 4.0)), gmin);
        hi2 = min(static_cast<long>(floor(p2 + 4.0)), gmax);
        da = static_cast<float>(static_cast<double>(lo1) - p1);
        db = static_cast<float>(static_cast<double>(lo2) - p2);
        f.q1 = -da;
        f.q2 = -db;
        const float w0 = in.f0[idx] ? 0.0f : (in.w0 ? in.w0[idx] : 1.0f);
        const float w3 = in.f3[idx] ? 0.0f : (in.w3 ? in.w3[idx] : 1.0f);
        const float ws = w0 + w3;
        if (ws > 0.0f) {
            v.x = (in.v0r[idx] * w0 + in.v3r[idx] * w3) / ws;
            v.y = (in.v0i[idx] * w0 + in.v3i[idx] * w3) / ws;
        } else {
            return false;                       // collapsed visibility is exactly zero
        }
    } else {
        const float p1 = in.bin[3 * idx] * gp.scale_f32;
        const float p2 = in.bin[3 * idx + 1] * gp.scale_f32;
        const float r3 = in.bin[3 * idx + 2];
        float centre;
        if (gp.archive) {
            centre = gp.r3_centre;
        } else {
            slab = static_cast<int>(floorf((r3 - gp.r3_min) / gp.slab_width));
            slab = min(max(slab, 0), gp.total_slabs - 1);
            if (slab < gp.slab_start || slab >= gp.slab_start + gp.nslabs) return false;
            centre = gp.r3_min + (static_cast<float>(slab) + 0.5f) * gp.slab_width;
        }
        s = fminf(1.0f, fmaxf(-1.0f, (r3 - centre) * gp.inv_half));
        lo1 = max(static_cast<long>(ceilf(p1 - 4)), gmin);
        hi1 = min(static_cast<long>(floorf(p1 + 4)), gmax);
        lo2 = max(static_cast<long>(ceilf(p2 - 4)), gmin);
        hi2 = min(static_cast<long>(floorf(p2 + 4)), gmax);
        da = static_cast<float>(lo1) - p1;
        db = static_cast<float>(lo2) - p2;
        f.q1 = p1;
        f.q2 = p2;
        v.x = in.vis_re[idx];
        v.y = in.vis_im[idx];
        if (v.x == 0.0f && v.y == 0.0f) return false;
    }
    // Non-finite coordinates convert to arbitrary integers: reject them, and any footprint
    // wider than the 9 cells a finite position can cover.
    if (!(fabsf(da) <= 5.0f) || !(fabsf(db) <= 5.0f)) return false;
    if (lo1 > hi1 || lo2 > hi2 || hi1 - lo1 > 8 || hi2 - lo2 > 8) return false;
    f.a0 = static_cast<int>(lo1 + off);
    f.na = static_cast<int>(hi1 - lo1 + 1);
    f.b0 = static_cast<int>(lo2 + off);
    f.nb = static_cast<int>(hi2 - lo2 + 1);
    f.da = da;
    f.db = db;
    f.local_slab = slab - (gp.archive ? 0 : gp.slab_start);
    f.s = s;
    f.vw.x = v.x / gp.det;
    f.vw.y = v.y / gp.det;
    return true;
}

}  // namespace svdi
