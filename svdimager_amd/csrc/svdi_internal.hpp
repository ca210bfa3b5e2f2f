// Internal declarations of libsvdimager: the per-handle context, the device-side geometry and
// plan records, and the stage launchers implemented in the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "svdi_math.hpp"
#include "svdimager.h"

namespace svdi {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define SVDI_HIP(call)                                                                   \
    do {                                                                                 \
        const hipError_t e__ = (call);                                                   \
        if (e__ != hipSuccess)                                                           \
            throw ::svdi::HipError(e__, std::string(#call) + ": " + hipGetErrorString(e__)); \
    } while (0)

#define SVDI_LAUNCH_CHECK() SVDI_HIP(hipGetLastError())

// Rotation of the snapshot, written on the device by the preprocess finalise kernel.
struct DevGeometry {
    double R[9];          // rows = principal axes, oriented (FP64)
    float vin[9];         // the same rows in FP32 (the reference's Vin)
    double evals[3];      // ascending
    double mean[3];       // mean baseline in wavelengths
    double fmin, fmax;    // channel frequency range, Hz
    float det;            // |V00 V11 - V01 V10|
};

// Chebyshev / slab plan, filled on the device, read back by the host in one copy.
struct DevPlan {
    float r3_min, r3_max, max_abs_t3;
    uint32_t valid_pixels;
    float r3_centre, r3_half_range, alpha_max, target_alpha;
    int32_t num_slabs;
    float r3_range, slab_width, slab_half_range, inv_slab_half_range, slab_alpha_max;
    uint32_t bucket_err_bits[4];     // integrated-imager buckets K = 8, 16, 24, 32
    uint32_t term_err_bits[32];      // archive search, K = 1 .. 32
    uint32_t vis_absmax_bits;        // max |vis| / |det| over gridded samples (for MFMA scaling)
    float vin[9];                    // the rotation rows the plan was made with (read back with it)
    // Values the host derives from the plan, formed the same way on the device: 1 / slab width,
    // |det| of the first two rotation rows (the vis scale) and its reciprocal.
    double inv_slab_width;
    float det, inv_det;
};

enum class Semantics { Integrated = SVDI_SEMANTICS_INTEGRATED, Archive = SVDI_SEMANTICS_ARCHIVE };

// Per-sample input of the imager: either the reference's preprocessed Bin / Vin / vis
// (FIpipe and FIpipeDevice) or a raw snapshot that is rotated and collapsed on the fly.
struct SampleInput {
    bool from_snapshot = false;
    size_t num_samples = 0;
    // preprocessed form
    const float* bin = nullptr;      // S x 3
    const float* vis_re = nullptr;
    const float* vis_im = nullptr;
    const float* vin_dev = nullptr;  // 9 floats on the device
    // snapshot form
    size_t num_rows = 0, num_channels = 0;
    const double* rot = nullptr;     // R x 3 rotated uvw (metres), device
    const double* freq = nullptr;
    const float *v0r = nullptr, *v0i = nullptr, *v3r = nullptr, *v3i = nullptr;
    const uint8_t *f0 = nullptr, *f3 = nullptr;
    const float *w0 = nullptr, *w3 = nullptr;
    // Rows a multiple of 4 and the per-sample arrays aligned for 4-sample loads (16 bytes for the
    // FP32 arrays, 4 for the flags): the binning passes then read four rows per load.
    bool quad_rows = false;
    const DevGeometry* geo = nullptr;
    // r3 range partials (min, max pairs) already reduced per block, or null to reduce Bin
    const float* r3_partials = nullptr;
    int r3_blocks = 0;
};

// Host copy of the plan plus the derived launch sizes.
struct HostPlan {
    DevPlan dev;
    Semantics semantics;
    int num_terms = 0;
    float term_error = 0.0f;
    size_t image_size = 0, grid_size = 0;
    float cell = 0.0f;
    size_t batch_slabs = 0;
    // FFT windows: the image pass reads rows [row_lo, row_lo + rows) and columns
    // [col_lo, col_lo + cols) of each transformed plane.
    size_t row_lo = 0, rows = 0, col_lo = 0, cols = 0;
};

struct GridParams {
    int G;
    int K;
    int slab_start, nslabs, total_slabs;
    int archive;            // 1: one slab centred at r3_centre
    float r3_min, slab_width, inv_half, r3_centre;
    float scale_f32;        // cell * G in FP32 (preprocessed input)
    double scale_f64;       // cell * G / c (snapshot input)
    float inv_det;          // 1 / |det| (vis weighting)
    float det;
    double inv_slab_width;  // 1 / slab_width
    int prec;               // MFMA gridder operand precision: 0 split3, 1 split c, 2 single
    // MFMA gridder work shape (svdi_tuning; 0 = the default).
    uint32_t coop_min, coop_cap;
    int class_log2;
    // Non-null for a call enqueued before its plan was read back: the kernels then take the plan's
    // values (r3 range, slab width, centre, vis scale) from this device record instead.
    const DevPlan* dplan;
};

// Grown-on-demand device buffers owned by a context.  In arena mode (between begin_arena and
// end_arena) the slots are carved instead, in request order, from a caller-given device range,
// and nothing is allocated or freed; a request beyond the range throws ArgError.
class Workspace {
  public:
    enum Slot {
        kPartials, kPlan, kGeometry, kRot, kConv, kPlanes, kRowFft, kColFft, kTwiddles,
        kBinCounts, kBinOffsets, kBinCursor, kRecPos, kRecVim, kItems, kMisc, kImage,
        kInBin, kInVre, kInVim, kInVin, kInSnap, kR3Partials, kScanSums, kRunPos, kRunOff, kItemCounts, kItemOffsets, kWinPartial, kSlots
    };
    void* get(Slot s, size_t bytes);
    void release();
    size_t bytes_held() const;
    void begin_arena(void* base, size_t capacity);
    void end_arena();
    // Frozen: a slot that would have to grow throws ArgError instead of allocating.
    void set_frozen(bool on) { frozen_ = on; }
    ~Workspace() { release(); }

  private:
    void* ptr_[kSlots] = {};
    size_t size_[kSlots] = {};
    // arena mode: the carved range, and the allocated slots put aside meanwhile
    bool frozen_ = false;
    char* arena_ = nullptr;
    size_t arena_cap_ = 0, arena_used_ = 0;
    void* kept_ptr_[kSlots] = {};
    size_t kept_size_[kSlots] = {};
};

}  // namespace svdi

namespace svdi {
// The sizes and options that decide a call's plan shape (its K and slab count once the geometry
// is known): a context keeps the K of the last one-slab call per key, the guess on which a
// snapshot call is enqueued before its plan is read back.
struct ShapeKey {
    size_t N;
    uint32_t cell_bits, alpha_bits;
    int num_terms, gridder, semantics;
    bool operator<(const ShapeKey& o) const {
        return std::tie(N, cell_bits, alpha_bits, num_terms, gridder, semantics) <
               std::tie(o.N, o.cell_bits, o.alpha_bits, o.num_terms, o.gridder, o.semantics);
    }
};

// What the host reads back, in one page-locked block: the plan (with its rotation rows, whose
// first two give the vis scale |det|) between planning and gridding, the gridder's totals and the
// consistency word at the end of a call.
struct PinnedPlan {
    DevPlan plan;
    uint32_t totals[8];     // the last slab batch's gridder totals ([0] = records)
    uint32_t error_word;    // copy of the context's device error word (nonzero: binning mismatch)
};

// Bits of the device error word that the binning kernels set.
constexpr uint32_t kErrBinMismatch = 1u;   // a block's scatter emitted other than it counted
}  // namespace svdi

struct svdi_context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_plan = nullptr;
    svdi::Workspace ws;
    svdi::PinnedPlan* pinned = nullptr;      // page-locked read-back of the plan
    // Device word the binning kernels OR kErr* bits into; cleared after each read-back.
    uint32_t* dev_err = nullptr;
    // This device's no-allocation scratch (the address of the library's static array there).
    void* noalloc_scratch = nullptr;
    // Batched calls: child contexts, each with its own stream and workspace, that image
    // snapshots concurrently.
    std::vector<std::unique_ptr<svdi_context>> lanes;
    std::mutex mu;
    size_t twiddle_n = 0;                    // FFT twiddles held for G = twiddle_n at twiddle_at
    const void* twiddle_at = nullptr;
    // Grid-correction table held in the workspace: computed for (N, G) at this address.
    size_t conv_n = 0, conv_g = 0;
    const float* conv_at = nullptr;
    int cu_count = 0;
    // K of the last one-slab snapshot call per shape key (the guess of a speculative call).
    std::map<svdi::ShapeKey, int> shapes;
    // CUs the window kernel's persistent launch leaves free: nonzero on the lanes of a batched
    // call, whose other lanes' kernels (binning, FFT, image) then run beside it.
    int grid_reserve = 0;
    // stage timing: pairs of events recorded around stages of the current call
    bool profiling = false;
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
    std::vector<std::pair<int, size_t>> stage_marks;   // (stage, index of start event)
    float stage_ms[SVDI_NUM_STAGES] = {};
    // MFMA gridder work shapes by grid size (0: every size without its own entry).
    std::map<size_t, svdi_tuning> tuning;
    svdi_context() = default;
    svdi_context(const svdi_context&) = delete;
    svdi_context& operator=(const svdi_context&) = delete;
    ~svdi_context();                         // waits for the context's work, frees everything
};

namespace svdi {
// Records a start event for `stage` now and returns a token; stage_end records its end.
size_t stage_begin(svdi_context* ctx, int stage);
void stage_end(svdi_context* ctx, size_t token);
// Raise a kernel's dynamic-LDS limit on the current device to at least `bytes`; the runtime is
// only called when the kernel needs more than it was last given there.
void allow_dynamic_lds(const void* kernel, size_t bytes);
}  // namespace svdi

namespace svdi {

// preprocess.hip
void launch_preprocess_geometry(svdi_context* ctx, const svdi_snapshot& snap, DevGeometry* geo,
                                double* rot, float* r3_partials, int* r3_blocks);
void launch_materialize(svdi_context* ctx, const svdi_snapshot& snap, const DevGeometry* geo,
                        float* bin, float* vin, float* vis_re, float* vis_im);

// plan.hip
void launch_plan(svdi_context* ctx, const SampleInput& in, Semantics sem, size_t image_size,
                 float cell, float target_alpha, DevPlan* plan);

// grid.hip
void launch_zero(svdi_context* ctx, void* p, size_t bytes);
void launch_fill(svdi_context* ctx, float* p, size_t count, float value);
void launch_grid_atomic(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                        float2* planes);
// tiles.hip (FP32 tiles) and windows.hip (MFMA windows): return a device pointer to
// {records, items} of the launch.
const uint32_t* launch_grid_tiled(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                                  float2* planes);
const uint32_t* launch_grid_windows(svdi_context* ctx, const SampleInput& in, const GridParams& gp,
                                    float2* planes);
// The window gridder in three steps, so that one contraction launch (on the calling context's
// stream) can take several slab batches, of several snapshots, binned on other streams:
// bin_windows enqueues the binning of one slab batch, contract_windows the contraction of up to
// kMaxWindowJobs binned batches of one MFMA precision, reduce_windows the sum of one batch's
// split windows; launch_grid_windows is the three for one batch.
constexpr int kMaxWindowJobs = 16;
struct WindowJob {
    const void* items = nullptr;
    uint32_t* totals = nullptr;            // device: [0] records, [1] wavefront items, ...
    void* records = nullptr;
    float2* planes = nullptr;
    float2* partial = nullptr;
    float2* trash = nullptr;
    const uint32_t* boff = nullptr;
    const uint32_t* ioff = nullptr;
    const uint32_t* multi = nullptr;
    uint32_t max_slots = 0;
    int K = 0, G = 0, prec = 0;
    int geom[16] = {};                     // the window geometry (windows.hip)
};
WindowJob bin_windows(svdi_context* ctx, const SampleInput& in, const GridParams& gp, float2* planes);
void contract_windows(svdi_context* ctx, const WindowJob* jobs, int count);
void reduce_windows(svdi_context* ctx, const WindowJob& job);
// api.cpp: the batched call (status), and svdi_last_error's message for failures reported
// elsewhere.
int image_batch(svdi_context* ctx, const svdi_snapshot* snaps, size_t count, float* const* images,
                const size_t* sizes, const float* cells, const svdi_options* opts, svdi_plan_info* infos,
                bool wait);
void report_error(const char* msg);
// Whether the window gridder takes this grid and sample count (else launch_grid_windows throws
// UnsupportedError).
bool grid_windows_supported(const SampleInput& in, const GridParams& gp);

// fft.hip
void fft_prepare(svdi_context* ctx, size_t G);
void launch_fft2_inverse(svdi_context* ctx, const float2* planes, size_t P, size_t G,
                         size_t row_lo, size_t rows, size_t col_lo, size_t cols,
                         float2* row_out, float2* out);

// image.hip
void launch_conv_corr(svdi_context* ctx, size_t N, size_t G, int archive_variant, float* conv);
struct ImageParams {
    int N, G;
    int P;                  // planes in this batch
    int K;
    int slab_start, nslabs;
    int archive;
    float cell;
    float r3_min, slab_width, slab_half, r3_centre;
    float inv_num_samples;
    int rows, cols;         // window sizes of the transformed planes
    int row_lo, col_lo;
    int signed_out;
    int lin_q0, lin_r0;     // (G G / 2 + G / 2) = lin_q0 G + lin_r0 (set by launch_image)
    const DevPlan* dplan;   // as GridParams::dplan
};

// The plan's values from the device record, for a call enqueued before its plan was read back
// (the same values, formed the same way, as the host copies into the parameters otherwise).
__device__ inline void take_device_plan(GridParams& gp) {
    const DevPlan* p = gp.dplan;
    if (p == nullptr) return;
    gp.r3_min = p->r3_min;
    gp.slab_width = p->slab_width;
    gp.inv_slab_width = p->inv_slab_width;
    gp.inv_half = p->inv_slab_half_range;
    gp.r3_centre = p->r3_centre;
    gp.det = p->det;
    gp.inv_det = p->inv_det;
}
__device__ inline void take_device_plan(ImageParams& ip) {
    const DevPlan* p = ip.dplan;
    if (p == nullptr) return;
    ip.r3_min = p->r3_min;
    ip.slab_width = p->slab_width;
    ip.slab_half = p->slab_half_range;
    ip.r3_centre = p->r3_centre;
}
void launch_image(svdi_context* ctx, const float2* planes, const float* vin_dev,
                  const float* conv, const ImageParams& ip, float* image);

}  // namespace svdi
