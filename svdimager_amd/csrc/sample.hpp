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

// Footprint of a sample at grid position (p1, p2) (centred coordinates, cells): the cells
// ceil(p - 4) .. floor(p + 4) clipped to [-G/2, (G-1)/2] on each axis (SVDIntegratedImager.cu:
// 1261-1275).  Non-finite positions convert to arbitrary integers: they are rejected, and so is
// any footprint wider than the 9 cells a finite position can cover.
__device__ __forceinline__ bool footprint_at(const GridParams& gp, double p1, double p2, Footprint& f) {
    const int gmin = -gp.G / 2;
    const int gmax = (gp.G - 1) / 2;
    const int off = gp.G / 2;
    // Positions clamped to +-1e9 cells (a NaN to -1e9) so that the conversions to int are defined;
    // such a position has no footprint cell inside the grid.
    const double c1 = fmin(fmax(p1, -1.0e9), 1.0e9), c2 = fmin(fmax(p2, -1.0e9), 1.0e9);
    const int lo1 = max(static_cast<int>(ceil(c1 - 4.0)), gmin);
    const int hi1 = min(static_cast<int>(floor(c1 + 4.0)), gmax);
    const int lo2 = max(static_cast<int>(ceil(c2 - 4.0)), gmin);
    const int hi2 = min(static_cast<int>(floor(c2 + 4.0)), gmax);
    const float da = static_cast<float>(static_cast<double>(lo1) - p1);
    const float db = static_cast<float>(static_cast<double>(lo2) - p2);
    if (!(fabsf(da) <= 5.0f) || !(fabsf(db) <= 5.0f)) return false;
    if (lo1 > hi1 || lo2 > hi2 || hi1 - lo1 > 8 || hi2 - lo2 > 8) return false;
    f.a0 = lo1 + off;
    f.na = hi1 - lo1 + 1;
    f.b0 = lo2 + off;
    f.nb = hi2 - lo2 + 1;
    f.da = da;
    f.db = db;
    return true;
}

// Snapshot input, one sample in two steps: the loads (load_snapshot_sample), then the
// arithmetic on them (finish_snapshot_sample), so that a thread can issue the loads of several
// samples before it needs the first.  (ru, rv, rw) is the sample's rotated row in metres.
struct SnapLoad {
    double nu;              // channel frequency, Hz
    float v0r, v0i, v3r, v3i, w0, w3;
    uint8_t f0, f3;
};

// kVis = false leaves out the visibilities (the fields are then zero); kWeights = false leaves
// out the weights (then 1), for callers whose decisions depend on flags and geometry only.
template <bool kVis = true, bool kWeights = true>
__device__ __forceinline__ SnapLoad load_snapshot_sample(const SampleInput& in, size_t idx, size_t chan) {
    SnapLoad s;
    s.nu = in.freq[chan];
    if constexpr (kVis) {
        s.v0r = in.v0r[idx];
        s.v0i = in.v0i[idx];
        s.v3r = in.v3r[idx];
        s.v3i = in.v3i[idx];
    } else {
        s.v0r = s.v0i = s.v3r = s.v3i = 0.0f;
    }
    s.f0 = in.f0[idx];
    s.f3 = in.f3[idx];
    s.w0 = kWeights && in.w0 ? in.w0[idx] : 1.0f;
    s.w3 = kWeights && in.w3 ? in.w3[idx] : 1.0f;
    return s;
}

// Pol collapse (SVDIntegratedImager.cu:588-617): wp = flag_p ? 0 : weight_p, V = (w0 V0 +
// w3 V3) / (w0 + w3) if w0 + w3 > 0, else 0.  A sample with V = 0 adds nothing: it is skipped,
// or with kFlagsOnly (whether a sample is kept then depends on its flags alone, so a pass that
// does not load the weights keeps the same samples) it is only skipped when both polarisations
// are flagged and is otherwise kept with V = 0.  Position b = uvw nu / c in FP64, in cells; slab
// and Chebyshev argument from r3 (:1249-1254).
template <bool kFlagsOnly = false>
__device__ __forceinline__ bool finish_snapshot_sample(const GridParams& gp, double ru, double rv, double rw,
                                                       const SnapLoad& s, Footprint& f) {
    const float w0 = s.f0 ? 0.0f : s.w0;
    const float w3 = s.f3 ? 0.0f : s.w3;
    const float ws = w0 + w3;
    if constexpr (kFlagsOnly) {
        if (s.f0 && s.f3) return false;
    } else {
        if (!(ws > 0.0f)) return false;
    }
    const double sc = s.nu * (1.0 / kSpeedOfLight);
    const double p1 = ru * sc * gp.scale_f64;
    const double p2 = rv * sc * gp.scale_f64;
    const double r3 = rw * sc;
    double centre;
    int slab = 0;
    if (gp.archive) {
        centre = gp.r3_centre;
    } else {
        if (gp.total_slabs > 1) {
            slab = static_cast<int>(floor((r3 - static_cast<double>(gp.r3_min)) * gp.inv_slab_width));
            slab = min(max(slab, 0), gp.total_slabs - 1);
        }
        if (slab < gp.slab_start || slab >= gp.slab_start + gp.nslabs) return false;
        centre = gp.r3_min + (static_cast<float>(slab) + 0.5f) * gp.slab_width;
    }
    if (!footprint_at(gp, p1, p2, f)) return false;
    f.q1 = -f.da;
    f.q2 = -f.db;
    f.local_slab = slab - (gp.archive ? 0 : gp.slab_start);
    f.s = static_cast<float>(fmin(1.0, fmax(-1.0, (r3 - centre) * static_cast<double>(gp.inv_half))));
    if (ws > 0.0f) {
        const float scale = gp.inv_det / ws;
        f.vw.x = (s.v0r * w0 + s.v3r * w3) * scale;
        f.vw.y = (s.v0i * w0 + s.v3i * w3) * scale;
    } else {
        f.vw = make_float2(0.0f, 0.0f);
    }
    return true;
}

template <bool kSnap>
__device__ __forceinline__ bool setup_sample(const SampleInput& in, const GridParams& gp,
                                             size_t idx, size_t row, size_t chan, Footprint& f) {
    if constexpr (kSnap) {
        return finish_snapshot_sample(gp, in.rot[3 * row], in.rot[3 * row + 1], in.rot[3 * row + 2],
                                      load_snapshot_sample(in, idx, chan), f);
    } else {
        const float p1 = in.bin[3 * idx] * gp.scale_f32;
        const float p2 = in.bin[3 * idx + 1] * gp.scale_f32;
        const float r3 = in.bin[3 * idx + 2];
        float centre;
        int slab = 0;
        if (gp.archive) {
            centre = gp.r3_centre;
        } else {
            slab = static_cast<int>(floorf((r3 - gp.r3_min) / gp.slab_width));
            slab = min(max(slab, 0), gp.total_slabs - 1);
            if (slab < gp.slab_start || slab >= gp.slab_start + gp.nslabs) return false;
            centre = gp.r3_min + (static_cast<float>(slab) + 0.5f) * gp.slab_width;
        }
        const float2 v = make_float2(in.vis_re[idx], in.vis_im[idx]);
        if (v.x == 0.0f && v.y == 0.0f) return false;
        if (!footprint_at(gp, p1, p2, f)) return false;
        f.q1 = p1;
        f.q2 = p2;
        f.local_slab = slab - (gp.archive ? 0 : gp.slab_start);
        f.s = fminf(1.0f, fmaxf(-1.0f, (r3 - centre) * gp.inv_half));
        f.vw.x = v.x / gp.det;
        f.vw.y = v.y / gp.det;
        return true;
    }
}

}  // namespace svdi
