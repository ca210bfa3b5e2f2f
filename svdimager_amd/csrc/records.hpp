// Window-gridder records: 16 bytes per (sample, window column), written by the binning scatter
// and read by the MFMA window kernel (one 16-byte access each way).  The same source compiles for
// the device and for CPU probes.
//
// Word layout (bit 31 on the left):
//   w0 = Re v mantissa (signed, 24) | s bits 0-7
//   w1 = Im v mantissa (signed, 24) | s bits 8-15
//   w2 = row cell (signed, 6) | row fraction (18) | s bits 16-23
//   w3 = column cell (signed, 6) | column fraction (18) | v exponent (signed, 8)
// Visibility: both components share the exponent E of the larger, |v| < 2^E, and keep 24-bit
// mantissas (error <= 2^-23 of the larger component).  s in [-1, 1] is stored as
// rint(2^23 s) + 2^23 (24 bits, the top code standing for s = 1 too): error <= 2^-24, and 2^-23
// at s = 1; the decode s = 2^-23 code - 1 is exact.  A position p (cells, relative to the record's
// origin) is stored as its footprint's first cell c = ceil(p - 4) and fraction F = c + 4 - p in
// [0, 1), F to 2^-18 (error <= 1.9e-6 cells); F = 0 marks an integer p, whose footprint has a
// ninth cell.
#pragma once

#include <cmath>
#include <cstdint>

#include "svdi_math.hpp"

namespace svdi {

constexpr int kRecFracBits = 18;
constexpr float kRecFracScale = 262144.0f;                    // 2^18
constexpr float kRecFracInv = 1.0f / 262144.0f;
constexpr float kRecSScale = 1.0f / 8388608.0f;              // 2^-23
constexpr int kRecCellMin = -32, kRecCellMax = 31;

struct RecordWords {
    uint32_t w[4];
};

// Position p = base + q (base an integer shift, q in cells) -> (cell, fraction) bits in 31..8:
// the first footprint cell k = ceil(q - 4) = ceil(q) - 4 and F = ceil(q) - q to 2^-18, formed
// exactly in FP32 and integers: rint(F 2^18) = ceil(q) 2^18 - rint(q 2^18), since ceil(q) 2^18 is
// an even integer (round-half-even commutes with shifts by even integers) and q 2^18 is exact.
// q is first clamped to [-8000, 8000] (NaN to -8000), which keeps ceil(q) 2^18 within int32; any
// |q| beyond that is far outside the cell range and clamps there anyway.
SVDI_HD uint32_t rec_encode_pos(int base, float q) {
    q = fminf(fmaxf(q, -8000.0f), 8000.0f);
    const int c = static_cast<int>(ceilf(q));
    int cell = base + c - 4;
    int frac = c * (1 << kRecFracBits) - static_cast<int>(rintf(q * kRecFracScale));
    if (frac >= (1 << kRecFracBits)) {              // F rounded to 1: the same point as (cell - 1, 0)
        frac -= 1 << kRecFracBits;
        cell -= 1;
    }
    cell = cell < kRecCellMin ? kRecCellMin : (cell > kRecCellMax ? kRecCellMax : cell);
    return (static_cast<uint32_t>(cell) << 26) | (static_cast<uint32_t>(frac) << 8);
}

SVDI_HD RecordWords rec_encode(int base1, float q1, int base2, float q2, float s, float vr, float vi) {
    // Shared exponent: |vr|, |vi| < 2^E; mantissas rint(v 2^(23 - E)) in [-2^23, 2^23 - 1].
    // |rint(v 2^(23 - E))| <= 2^24 (2^23 unless E was clamped), so the clamps and the
    // conversions are done on integer-valued floats, all exact.
    const float m = fmaxf(fabsf(vr), fabsf(vi));
    int e = 0;
    if (m > 0.0f) (void)frexpf(m, &e);
    e = e < -127 ? -127 : (e > 127 ? 127 : e);
    auto mant = [&](float v) {
        const float r = fminf(fmaxf(rintf(ldexpf(v, 23 - e)), -8388608.0f), 8388607.0f);
        return static_cast<uint32_t>(static_cast<int32_t>(r)) << 8;
    };
    const float sc = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);
    const float sq = fminf(fmaxf(rintf(sc * 8388608.0f) + 8388608.0f, 0.0f), 16777215.0f);
    const uint32_t su = static_cast<uint32_t>(sq);
    RecordWords r;
    r.w[0] = mant(vr) | (su & 0xffu);
    r.w[1] = mant(vi) | ((su >> 8) & 0xffu);
    r.w[2] = rec_encode_pos(base1, q1) | ((su >> 16) & 0xffu);
    r.w[3] = rec_encode_pos(base2, q2) | (static_cast<uint32_t>(e) & 0xffu);
    return r;
}

// Decoders (the window kernel inlines the same expressions).
SVDI_HD int rec_cell(uint32_t w) { return static_cast<int32_t>(w) >> 26; }
SVDI_HD uint32_t rec_frac_bits(uint32_t w) { return (w >> 8) & ((1u << kRecFracBits) - 1u); }
SVDI_HD float rec_s(uint32_t w0, uint32_t w1, uint32_t w2) {
    const uint32_t su = (w0 & 0xffu) | ((w1 & 0xffu) << 8) | ((w2 & 0xffu) << 16);
    return fmaf(static_cast<float>(su), kRecSScale, -1.0f);
}
SVDI_HD int rec_exponent(uint32_t w3) { return static_cast<int32_t>(w3 << 24) >> 24; }
SVDI_HD float rec_v(uint32_t w, uint32_t w3, int extra_exp) {
    return ldexpf(static_cast<float>(static_cast<int32_t>(w) >> 8), rec_exponent(w3) - 23 + extra_exp);
}
// The position itself (p = cell + 4 - F), for tests.
SVDI_HD float rec_pos(uint32_t w) {
    return static_cast<float>(rec_cell(w) + 4) - static_cast<float>(rec_frac_bits(w)) * kRecFracInv;
}

}  // namespace svdi
