// TIME snapshots from exported MeasurementSet columns (msreader.cpp): host-only.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace svdi {

// HostMeasurementSetData + MeasurementSetSnapshot (SVDIntegratedImager.cu:39-61), with uvw and
// the channel frequencies in FP64.  Sample index = row + chan * num_rows.
struct MsSnapshot {
    size_t index = 0, total_snapshots = 0;
    double time = 0.0;
    size_t num_rows = 0, num_channels = 0;
    std::vector<double> uvw, freq_hz;
    std::vector<float> vis0_re, vis0_im, vis3_re, vis3_im, weight0, weight3;
    std::vector<uint8_t> flag0, flag3;
    bool has_weights = false;
    double array_center_x = 0.0;
};

// Snapshot `snapshot_index` (TIME order) of the column directory `dir`; throws
// std::runtime_error with the reference's messages where it has them.
MsSnapshot read_ms_snapshot(const std::string& dir, size_t snapshot_index);

}  // namespace svdi
