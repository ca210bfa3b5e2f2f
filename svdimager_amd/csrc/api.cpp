// C-ABI of libsvdimager and the host orchestration of one imaging call.
//
// Sequencing follows FIpipeDevice (SVDIntegratedImager.cu:1498-1761): plan (r3 range, t3
// stats, slab rule, Chebyshev K) -> per slab batch { grid -> 2D inverse FFT -> image pass }.
// Differences: the plan is finalised on the device and read back in one copy; buffers, the
// stream and FFT twiddles live in the context and are reused across calls; errors are status
// codes, never exceptions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>

#include "msreader.hpp"
#include "svdi_internal.hpp"

// Device scratch of the no-allocation entry point (svdi_image_snapshot_noalloc): a file-scope
// array of which every GPU the library runs on holds its own copy until the process ends.  Host
// code names it only in hipGetSymbolAddress, once per device at context creation; everywhere
// else (kernel arguments, fills) the address returned there is used.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace svdi {

namespace {
thread_local std::string g_last_error;

// The device address of `scratch` on each device, taken with that device current.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_at;

void* noalloc_scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_at.size() <= static_cast<size_t>(device)) g_scratch_at.resize(device + 1, nullptr);
    if (g_scratch_at[device] == nullptr) {
        void* p = nullptr;
        SVDI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_at[device] = p;
    }
    return g_scratch_at[device];
}

void set_error(const std::string& s) { g_last_error = s; }

uint32_t bits_of(float f) {
    uint32_t b;
    std::memcpy(&b, &f, sizeof b);
    return b;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return SVDI_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SVDI_ERR_INVALID_ARGUMENT;
    } catch (const UnsupportedError& e) {
        set_error(e.what());
        return SVDI_ERR_UNSUPPORTED;
    } catch (const HipError& e) {
        set_error(e.what());
        return e.code == hipErrorOutOfMemory ? SVDI_ERR_OUT_OF_MEMORY : SVDI_ERR_HIP;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return SVDI_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SVDI_ERR_INTERNAL;
    } catch (...) {
        set_error("unknown error");
        return SVDI_ERR_INTERNAL;
    }
}

// SVD_CHEB_TARGET_ALPHA, ignored unless positive (SVDIntegratedImager.cu:963-971).
float env_target_alpha() {
    if (const char* s = std::getenv("SVD_CHEB_TARGET_ALPHA")) {
        const float v = static_cast<float>(std::atof(s));
        if (v > 0.0f) return v;
    }
    return kDefaultTargetAlpha;
}

// choose_slab_batch_size (SVDIntegratedImager.cu:1470-1496).
size_t choose_batch_slabs(size_t total_slabs, int K, size_t plane_elems, size_t opt_max) {
    if (opt_max > 0) return std::min(total_slabs, opt_max);
    const char* env = std::getenv("SVDIMAGER_MAX_BATCH_SLABS");
    if (env != nullptr && env[0] != '\0') {
        size_t requested = 0;
        try {
            requested = std::stoull(env);
        } catch (const std::exception&) {
            throw ArgError("SVDIMAGER_MAX_BATCH_SLABS is not an unsigned integer");
        }
        if (requested == 0) throw ArgError("SVDIMAGER_MAX_BATCH_SLABS must be greater than zero");
        return std::min(total_slabs, requested);
    }
    size_t free_b = 0, total_b = 0;
    SVDI_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_slab = static_cast<size_t>(K) * plane_elems * sizeof(float2);
    const size_t budget = free_b * 25 / 100;
    return std::min(total_slabs, std::max<size_t>(1, budget / per_slab));
}

float bits_to_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, sizeof f);
    return f;
}

void print_report(const HostPlan& hp, std::ostream& os) {
    const DevPlan& p = hp.dev;
    if (hp.semantics == Semantics::Integrated) {
        os << "Chebyshev summary for paper:\n";
        os << "  Phi_max                           = " << p.alpha_max << "\n";
        os << "  alpha_target                      = " << p.target_alpha << "\n";
        os << "  N_slab                            = " << p.num_slabs << "\n";
        os << "  alpha_slab                        = " << p.slab_alpha_max << "\n";
        os << "  selected K                        = " << hp.num_terms << "\n";
        os << "  N_slab * K moment reconstructions = "
           << static_cast<size_t>(p.num_slabs) * static_cast<size_t>(hp.num_terms) << "\n";
        os << "  target approximation error         = " << kTargetError << "\n";
    } else {
        os << "Chebyshev range check:\n";
        os << "  r3 interval = [" << p.r3_min << ", " << p.r3_max << "] wavelengths\n";
        os << "  r3 centre   = " << p.r3_centre << " wavelengths\n";
        os << "  r3 halfspan = " << p.r3_half_range << " wavelengths\n";
        os << "  max |t3|    = " << p.max_abs_t3 << " over " << p.valid_pixels << " valid pixels\n";
        os << "  alpha_max   = 2*pi*max|t3|*r3_halfspan = " << p.alpha_max << "\n";
        os << "  target sampled max error            = " << kTargetError << "\n";
        os << "  selected kept terms                 = " << hp.num_terms << " (keep T0..T"
           << (hp.num_terms - 1) << ")\n";
        os << "  selected sampled max error          = " << hp.term_error << "\n";
        if (hp.term_error < kTargetError) {
            os << "  Status: workable (sampled error < target)\n";
        } else if (hp.term_error < 5.0f * kTargetError) {
            os << "  Status: marginal (sampled error between target and 5x target)\n";
        } else {
            os << "  Status: not safe within configured max terms on this dataset\n";
        }
    }
}

void check_device_pointer(const void* p, const char* what) {
    if (p == nullptr) throw ArgError(std::string(what) + " is null");
}

void begin_call(svdi_context* ctx) {
    ctx->events_used = 0;
    ctx->stage_marks.clear();
    for (int i = 0; i < SVDI_NUM_STAGES; ++i) ctx->stage_ms[i] = 0.0f;
}

// After the stream has been synchronised: sum the recorded stage intervals.
void finish_call(svdi_context* ctx) {
    for (const auto& m : ctx->stage_marks) {
        float ms = 0.0f;
        SVDI_HIP(hipEventElapsedTime(&ms, ctx->event_pool[m.second], ctx->event_pool[m.second + 1]));
        ctx->stage_ms[m.first] += ms;
    }
}

// Concurrent imaging lanes of a batched call.
constexpr size_t kBatchLanes = 4;

// Host-side state of one imaging call between its phases.
struct CallState {
    HostPlan hp;
    size_t N = 0;
    float cell = 0.0f;
    size_t records = 0;
    const uint32_t* tile_totals = nullptr;
    int gridder = SVDI_GRIDDER_AUTO;                       // resolved by image_enqueue
    float det = 0.0f;                                      // |V00 V11 - V01 V10| of the rotation rows
    int early = 0;                                         // svdi_plan_info::enqueued_early
};

// Phase 1 of an imaging call: validate, enqueue the planning kernels and the read-back of the
// plan and the rotation rows into page-locked memory, and record ev_plan.  `in` describes the
// samples.
void image_begin(svdi_context* ctx, const SampleInput& in, size_t N, float cell, const svdi_options& o,
                 CallState& cs) {
    if (N == 0) throw ArgError("Image_Size must be greater than zero");
    if (!(cell > 0.0f)) throw ArgError("Cell_Size must be greater than zero");
    if (in.num_samples == 0) throw ArgError("Measurement set is empty");
    if (o.num_terms < 0 || o.num_terms > kMaxTerms) throw ArgError("num_terms must be in 0..32");
    if (o.mfma_precision < SVDI_PRECISION_DEFAULT || o.mfma_precision > SVDI_PRECISION_BF16)
        throw ArgError("unknown mfma_precision");
    if (N > 46340) throw ArgError("Image_Size too large");
    if (o.gridder != SVDI_GRIDDER_AUTO && o.gridder != SVDI_GRIDDER_ATOMIC && o.gridder != SVDI_GRIDDER_TILED &&
        o.gridder != SVDI_GRIDDER_MFMA)
        throw ArgError("unknown gridder");
    const size_t G = (N * 3 + 1) / 2;                      // ceiling_divide(image_size * 3, 2)
    cs = CallState();
    cs.N = N;
    cs.cell = cell;
    cs.hp.semantics = o.semantics == SVDI_SEMANTICS_ARCHIVE ? Semantics::Archive : Semantics::Integrated;
    cs.hp.image_size = N;
    cs.hp.grid_size = G;
    cs.hp.cell = cell;
    const float target_alpha = o.target_alpha > 0.0f ? o.target_alpha : env_target_alpha();

    DevPlan* dplan = static_cast<DevPlan*>(ctx->ws.get(Workspace::kPlan, sizeof(DevPlan)));
    SVDI_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    const size_t tok = stage_begin(ctx, 1);
    launch_plan(ctx, in, cs.hp.semantics, N, cell, target_alpha, dplan);
    stage_end(ctx, tok);
    SVDI_HIP(hipMemcpyAsync(&ctx->pinned->plan, dplan, sizeof(DevPlan), hipMemcpyDeviceToHost, ctx->stream));
    SVDI_HIP(hipEventRecord(ctx->ev_plan, ctx->stream));
}

// Phase 2a, once ev_plan has completed: take the plan and the rotation rows from the read-back,
// choose K and the slab batches on the host.
void image_take_plan(svdi_context* ctx, const svdi_options& o, CallState& cs) {
    HostPlan& hp = cs.hp;
    hp.dev = ctx->pinned->plan;
    cs.det = hp.dev.det;                                   // |v0 v4 - v1 v3| of the rotation rows

    // K: bucket rule (SVDIntegratedImager.cu:1577-1603) or smallest passing K (FIPkernels.cu:301-318).
    if (hp.semantics == Semantics::Integrated) {
        const int buckets[4] = {8, 16, 24, 32};
        int K = kMaxTerms;
        float err = bits_to_float(hp.dev.bucket_err_bits[3]);
        for (int i = 0; i < 4; ++i) {
            const float e = bits_to_float(hp.dev.bucket_err_bits[i]);
            if (e <= kTargetError) {
                K = buckets[i];
                err = e;
                break;
            }
        }
        hp.num_terms = K;
        hp.term_error = err;
    } else {
        int K = -1;
        for (int k = 1; k <= kMaxTerms; ++k)
            if (bits_to_float(hp.dev.term_err_bits[k - 1]) <= kTargetError) {
                K = k;
                break;
            }
        if (K < 0) K = kMaxTerms;
        hp.num_terms = K;
        hp.term_error = bits_to_float(hp.dev.term_err_bits[K - 1]);
    }
    if (o.num_terms > 0) {
        hp.num_terms = o.num_terms;
        hp.term_error = hp.semantics == Semantics::Archive ? bits_to_float(hp.dev.term_err_bits[o.num_terms - 1])
                                                           : hp.term_error;
    }
    hp.batch_slabs = choose_batch_slabs(static_cast<size_t>(hp.dev.num_slabs), hp.num_terms,
                                        hp.grid_size * hp.grid_size, o.max_batch_slabs);
}

// The MFMA gridder's work shape for grid size G: the context's entry for G, else its entry for
// every size, else the defaults (all zero).
svdi_tuning tuning_for(const svdi_context* ctx, size_t G) {
    auto it = ctx->tuning.find(G);
    if (it == ctx->tuning.end()) it = ctx->tuning.find(0);
    if (it != ctx->tuning.end()) return it->second;
    svdi_tuning t = {};
    return t;
}

// A call's buffers and parameters between the phases of image_enqueue.
struct Enqueued {
    GridParams gp;
    ImageParams ip;
    float2* planes = nullptr;
    float2* rowfft = nullptr;
    float2* colfft = nullptr;
    float* conv = nullptr;
    const float* vin_dev = nullptr;
    int gridder = SVDI_GRIDDER_AUTO;
    size_t total_slabs = 0;
};

// Phase 2b, with the plan in cs: the call's buffers, the image cleared, the grid-correction table,
// the gridding and image-pass parameters, and the gridder resolved.
Enqueued image_prepare(svdi_context* ctx, const SampleInput& in, const svdi_options& o, CallState& cs, float* image,
                       const DevPlan* dplan = nullptr) {
    Enqueued e;
    HostPlan& hp = cs.hp;
    const size_t N = cs.N, G = hp.grid_size;
    const float cell = cs.cell;
    const int K = hp.num_terms;
    const size_t total_slabs = static_cast<size_t>(hp.dev.num_slabs);
    e.total_slabs = total_slabs;

    if (G % 2 == 0) {
        hp.row_lo = G / 2 - N / 2;
        hp.col_lo = G / 2 - N / 2;
        hp.rows = hp.cols = N;
    } else {                                               // odd G: keep the whole plane
        hp.row_lo = hp.col_lo = 0;
        hp.rows = hp.cols = G;
    }
    const size_t P = hp.batch_slabs * static_cast<size_t>(K);
    e.planes = static_cast<float2*>(ctx->ws.get(Workspace::kPlanes, P * G * G * sizeof(float2)));
    e.rowfft = static_cast<float2*>(ctx->ws.get(Workspace::kRowFft, P * G * hp.cols * sizeof(float2)));
    e.colfft = static_cast<float2*>(ctx->ws.get(Workspace::kColFft, P * hp.rows * hp.cols * sizeof(float2)));
    float* conv = static_cast<float*>(ctx->ws.get(Workspace::kConv, (N / 2 + 1) * sizeof(float)));
    e.conv = conv;

    if (o.verbose) print_report(hp, std::cout);

    launch_fill(ctx, image, N * N, 0.0f);
    if (ctx->conv_at != conv || ctx->conv_n != N || ctx->conv_g != G) {   // depends on N and G only
        launch_conv_corr(ctx, N, G, hp.semantics == Semantics::Archive, conv);
        ctx->conv_at = conv;
        ctx->conv_n = N;
        ctx->conv_g = G;
    }

    GridParams& gp = e.gp;
    gp = GridParams();
    gp.G = static_cast<int>(G);
    gp.K = K;
    gp.total_slabs = static_cast<int>(total_slabs);
    gp.archive = hp.semantics == Semantics::Archive;
    gp.r3_min = hp.dev.r3_min;
    gp.slab_width = hp.dev.slab_width;
    gp.inv_slab_width = 1.0 / static_cast<double>(hp.dev.slab_width);
    gp.inv_half = hp.dev.inv_slab_half_range;
    gp.r3_centre = hp.dev.r3_centre;
    gp.scale_f32 = cell * static_cast<float>(G);
    gp.scale_f64 = static_cast<double>(cell) * static_cast<double>(G);
    gp.prec = o.mfma_precision == SVDI_PRECISION_SPLITC   ? 1
              : o.mfma_precision == SVDI_PRECISION_SINGLE ? 2
              : o.mfma_precision == SVDI_PRECISION_BF16   ? 3
                                                          : 0;
    gp.det = cs.det;
    gp.inv_det = 1.0f / cs.det;
    gp.dplan = dplan;
    const svdi_tuning tune = tuning_for(ctx, G);
    gp.coop_min = tune.coop_min;
    gp.coop_cap = tune.coop_cap;
    gp.class_log2 = tune.class_log2;

    ImageParams& ip = e.ip;
    ip = ImageParams();
    ip.N = static_cast<int>(N);
    ip.G = static_cast<int>(G);
    ip.K = K;
    ip.archive = gp.archive;
    ip.cell = cell;
    ip.r3_min = hp.dev.r3_min;
    ip.slab_width = hp.dev.slab_width;
    ip.slab_half = hp.dev.slab_half_range;
    ip.r3_centre = hp.dev.r3_centre;
    ip.inv_num_samples = 1.0f / static_cast<float>(in.num_samples);
    ip.rows = static_cast<int>(hp.rows);
    ip.cols = static_cast<int>(hp.cols);
    ip.row_lo = static_cast<int>(hp.row_lo);
    ip.col_lo = static_cast<int>(hp.col_lo);
    ip.signed_out = o.output == SVDI_OUTPUT_SIGNED;
    ip.dplan = dplan;
    e.vin_dev = in.from_snapshot ? in.geo->vin : in.vin_dev;

    // AUTO: the MFMA window gridder, or the FP32 tile gridder for the shapes it does not take
    // (judged at the largest slab batch).
    GridParams probe = gp;
    probe.nslabs = static_cast<int>(hp.batch_slabs);
    const int gridder = o.gridder != SVDI_GRIDDER_AUTO ? o.gridder
                        : grid_windows_supported(in, probe) ? SVDI_GRIDDER_MFMA : SVDI_GRIDDER_TILED;
    cs.gridder = gridder;
    e.gridder = gridder;
    return e;
}

// Phase 2c: the slab batch [start, start + nb) of a prepared call gridded on `ctx`'s stream.
void image_grid(svdi_context* ctx, const SampleInput& in, CallState& cs, Enqueued& e, size_t start, size_t nb) {
    e.gp.slab_start = static_cast<int>(start);
    e.gp.nslabs = static_cast<int>(nb);
    const size_t tok = stage_begin(ctx, 2);
    if (e.gridder == SVDI_GRIDDER_MFMA) {
        cs.tile_totals = launch_grid_windows(ctx, in, e.gp, e.planes);
    } else if (e.gridder == SVDI_GRIDDER_TILED) {
        cs.tile_totals = launch_grid_tiled(ctx, in, e.gp, e.planes);
    } else {
        launch_grid_atomic(ctx, in, e.gp, e.planes);
    }
    stage_end(ctx, tok);
}

// Phase 2d: the FFT and image pass of a gridded slab batch [start, start + nb).
void image_post(svdi_context* ctx, CallState& cs, Enqueued& e, size_t start, size_t nb, float* image) {
    const HostPlan& hp = cs.hp;
    const size_t Pb = nb * static_cast<size_t>(hp.num_terms);
    size_t tok = stage_begin(ctx, 3);
    launch_fft2_inverse(ctx, e.planes, Pb, hp.grid_size, hp.row_lo, hp.rows, hp.col_lo, hp.cols, e.rowfft, e.colfft);
    stage_end(ctx, tok);
    e.ip.P = static_cast<int>(Pb);
    e.ip.slab_start = static_cast<int>(start);
    e.ip.nslabs = static_cast<int>(nb);
    tok = stage_begin(ctx, 4);
    launch_image(ctx, e.colfft, e.vin_dev, e.conv, e.ip, image);
    stage_end(ctx, tok);
}

// Phase 2 with the plan in cs: gridding, FFT and image pass per slab batch.  The image (N x N
// floats, device) is overwritten.
void image_enqueue(svdi_context* ctx, const SampleInput& in, const svdi_options& o, CallState& cs, float* image,
                   const DevPlan* dplan = nullptr) {
    Enqueued e = image_prepare(ctx, in, o, cs, image, dplan);
    for (size_t start = 0; start < e.total_slabs; start += cs.hp.batch_slabs) {
        const size_t nb = std::min(cs.hp.batch_slabs, e.total_slabs - start);
        image_grid(ctx, in, cs, e, start, nb);
        image_post(ctx, cs, e, start, nb, image);
    }
}

// End of a call's work on `ctx`'s stream: the gridder totals (when `totals` is set) are copied
// into page-locked memory; with `read_error` the device error word is too, and is then cleared
// for the next call (without it the word keeps its bits until a later call reads it); ev_stop
// is recorded.  No null-stream operation is involved, so other contexts' work and the caller's
// streams are not waited for.
void image_finish_enqueue(svdi_context* ctx, const uint32_t* totals, bool read_error) {
    if (totals != nullptr)
        SVDI_HIP(hipMemcpyAsync(ctx->pinned->totals, totals, sizeof ctx->pinned->totals, hipMemcpyDeviceToHost,
                                ctx->stream));
    if (read_error) {
        SVDI_HIP(hipMemcpyAsync(&ctx->pinned->error_word, ctx->dev_err, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                ctx->stream));
        launch_fill(ctx, reinterpret_cast<float*>(ctx->dev_err), 1, 0.0f);
    }
    SVDI_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
}

// After ev_stop has completed: a binning inconsistency detected on the device becomes an error.
void check_error_word(const svdi_context* ctx) {
    if (ctx->pinned->error_word & kErrBinMismatch)
        throw std::runtime_error("gridder binning mismatch: a block's scatter pass emitted other records than "
                                 "its count pass counted (the inputs changed during the call?)");
}

void fill_info(const CallState& cs, float ms, svdi_plan_info* info) {
    if (!info) return;
    const HostPlan& hp = cs.hp;
    info->r3_min = hp.dev.r3_min;
    info->r3_max = hp.dev.r3_max;
    info->max_abs_t3 = hp.dev.max_abs_t3;
    info->valid_pixels = hp.dev.valid_pixels;
    info->alpha_max = hp.dev.alpha_max;
    info->target_alpha = hp.dev.target_alpha;
    info->num_slabs = hp.dev.num_slabs;
    info->slab_width = hp.dev.slab_width;
    info->slab_half_range = hp.dev.slab_half_range;
    info->slab_alpha_max = hp.dev.slab_alpha_max;
    info->num_terms = hp.num_terms;
    info->term_error = hp.term_error;
    for (int i = 0; i < 4; ++i) info->bucket_errors[i] = bits_to_float(hp.dev.bucket_err_bits[i]);
    info->batch_slabs = hp.batch_slabs;
    info->grid_size = hp.grid_size;
    info->num_planes = static_cast<size_t>(hp.dev.num_slabs) * static_cast<size_t>(hp.num_terms);
    info->records = cs.records;
    info->elapsed_ms = ms;
    info->gridder = cs.gridder;
    info->r3_centre = hp.dev.r3_centre;
    info->inv_slab_half_range = hp.dev.inv_slab_half_range;
    info->vis_det = cs.det;
    info->enqueued_early = cs.early;
}

// Phase 3: wait for the call's work, then report.
void image_complete(svdi_context* ctx, const svdi_options& o, CallState& cs, svdi_plan_info* info) {
    SVDI_HIP(hipEventSynchronize(ctx->ev_stop));
    finish_call(ctx);
    check_error_word(ctx);
    float ms = 0.0f;
    SVDI_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (cs.tile_totals != nullptr) cs.records = ctx->pinned->totals[0];   // last slab batch
    if (o.verbose) std::cout << "Time elapsed: " << ms << " ms" << std::endl;
    fill_info(cs, ms, info);
}

// Enqueueing a snapshot call before its plan is read back.  A context records, per image size,
// cell and plan-shaping options, the K of its last call that had one slab.  A snapshot call whose
// key has a record is enqueued whole right after its planning kernels, on the guess of one slab
// and that K, with the gridding and image kernels taking the plan's values from the device record.
// When the host has read the plan, it keeps the guessed work if the plan's shape is the guess, and
// otherwise enqueues the call's work again with the real plan behind it on the same stream, which
// overwrites every buffer and the image.  Either way the image is the one the plan calls for.
struct Guess {
    ShapeKey key{};
    int K = 0;                                             // 0: no guess
    CallState sp;                                          // the guessed call's state
};

// The guess for a call (K = 0 when there is none) and, with one, its work enqueued on `lane`.
Guess guess_and_enqueue(const std::map<ShapeKey, int>& shapes, svdi_context* lane, const SampleInput& in,
                        const svdi_options& o, const CallState& cs, float* image) {
    Guess g;
    g.key = ShapeKey{cs.N, bits_of(cs.cell), bits_of(o.target_alpha), o.num_terms, o.gridder, o.semantics};
    if (!in.from_snapshot || o.semantics != SVDI_SEMANTICS_INTEGRATED || o.verbose) return g;
    const auto it = shapes.find(g.key);
    if (it == shapes.end()) return g;
    g.K = it->second;
    g.sp = cs;
    g.sp.hp.num_terms = g.K;
    g.sp.hp.dev.num_slabs = 1;
    g.sp.hp.batch_slabs = 1;
    g.sp.det = 1.0f;                                       // (the kernels take the device plan's)
    const DevPlan* dplan = static_cast<const DevPlan*>(lane->ws.get(Workspace::kPlan, sizeof(DevPlan)));
    image_enqueue(lane, in, o, g.sp, image, dplan);
    return g;
}

// With the plan taken into cs: keep the guessed work or enqueue the call's work again, and record
// the call's shape.
void settle_guess(std::map<ShapeKey, int>& shapes, svdi_context* lane, const SampleInput& in, const svdi_options& o,
                  const Guess& g, CallState& cs, float* image) {
    const bool one_slab = cs.hp.dev.num_slabs == 1 && cs.hp.batch_slabs == 1;
    if (g.K > 0 && one_slab && cs.hp.num_terms == g.K) {
        cs.gridder = g.sp.gridder;
        cs.tile_totals = g.sp.tile_totals;
        cs.early = 1;
    } else {
        image_enqueue(lane, in, o, cs, image);
        cs.early = g.K > 0 ? 2 : 0;
    }
    if (one_slab) shapes[g.key] = cs.hp.num_terms;
    else shapes.erase(g.key);
}

// One imaging call on one context: the three phases back to back (with `speculate`, enqueued
// before the plan read-back when the context has a guess for it).
void image_core(svdi_context* ctx, const SampleInput& in, size_t N, float cell,
                const svdi_options& o, svdi_plan_info* info, float* image, bool speculate = false) {
    CallState cs;
    image_begin(ctx, in, N, cell, o, cs);
    if (speculate) {
        const Guess g = guess_and_enqueue(ctx->shapes, ctx, in, o, cs, image);
        SVDI_HIP(hipEventSynchronize(ctx->ev_plan));
        image_take_plan(ctx, o, cs);
        settle_guess(ctx->shapes, ctx, in, o, g, cs, image);
    } else {
        SVDI_HIP(hipEventSynchronize(ctx->ev_plan));
        image_take_plan(ctx, o, cs);
        image_enqueue(ctx, in, o, cs, image);
    }
    image_finish_enqueue(ctx, cs.tile_totals, true);
    image_complete(ctx, o, cs, info);
}

svdi_options resolve(const svdi_options* o, int default_semantics) {
    svdi_options r;
    svdi_default_options(&r);
    r.semantics = default_semantics;
    if (o) r = *o;
    return r;
}

struct DeviceGuard {
    explicit DeviceGuard(int dev) { SVDI_HIP(hipSetDevice(dev)); }
};

SampleInput snapshot_input(svdi_context* ctx, const svdi_snapshot& s) {
    if (s.num_rows == 0 || s.num_channels == 0) throw ArgError("Measurement set is empty");
    check_device_pointer(s.uvw, "uvw");
    check_device_pointer(s.freq_hz, "freq_hz");
    check_device_pointer(s.vis0_re, "vis0_re");
    check_device_pointer(s.vis0_im, "vis0_im");
    check_device_pointer(s.vis3_re, "vis3_re");
    check_device_pointer(s.vis3_im, "vis3_im");
    check_device_pointer(s.flag0, "flag0");
    check_device_pointer(s.flag3, "flag3");
    if (s.num_rows > 0xffffffffull || s.num_channels > 65535)
        throw ArgError("snapshot too large: rows must fit 32 bits and channels 65535");
    DevGeometry* geo = static_cast<DevGeometry*>(ctx->ws.get(Workspace::kGeometry, sizeof(DevGeometry)));
    double* rot = static_cast<double*>(ctx->ws.get(Workspace::kRot, s.num_rows * 3 * sizeof(double)));
    const size_t r3_blocks = (s.num_rows + 255) / 256;
    float* r3p = static_cast<float*>(ctx->ws.get(Workspace::kR3Partials, r3_blocks * 2 * sizeof(float)));
    int nb = 0;
    const size_t tok = stage_begin(ctx, 0);
    launch_preprocess_geometry(ctx, s, geo, rot, r3p, &nb);
    stage_end(ctx, tok);
    SampleInput in;
    in.from_snapshot = true;
    in.num_rows = s.num_rows;
    in.num_channels = s.num_channels;
    in.num_samples = s.num_rows * s.num_channels;
    in.rot = rot;
    in.freq = s.freq_hz;
    in.v0r = s.vis0_re;
    in.v0i = s.vis0_im;
    in.v3r = s.vis3_re;
    in.v3i = s.vis3_im;
    in.f0 = s.flag0;
    in.f3 = s.flag3;
    in.w0 = s.weight0;
    in.w3 = s.weight3;
    auto aligned = [](const void* p, uintptr_t a) { return p == nullptr || reinterpret_cast<uintptr_t>(p) % a == 0; };
    in.quad_rows = s.num_rows % 4 == 0 && aligned(s.vis0_re, 16) && aligned(s.vis0_im, 16) &&
                   aligned(s.vis3_re, 16) && aligned(s.vis3_im, 16) && aligned(s.weight0, 16) &&
                   aligned(s.weight3, 16) && aligned(s.flag0, 4) && aligned(s.flag3, 4);
    in.geo = geo;
    in.r3_partials = r3p;
    in.r3_blocks = nb;
    return in;
}

}  // namespace

size_t stage_begin(svdi_context* ctx, int stage) {
    if (!ctx->profiling) return SIZE_MAX;
    while (ctx->event_pool.size() < ctx->events_used + 2) {
        hipEvent_t e;
        SVDI_HIP(hipEventCreate(&e));
        ctx->event_pool.push_back(e);
    }
    const size_t idx = ctx->events_used;
    ctx->events_used += 2;
    SVDI_HIP(hipEventRecord(ctx->event_pool[idx], ctx->stream));
    ctx->stage_marks.emplace_back(stage, idx);
    return idx;
}

void stage_end(svdi_context* ctx, size_t token) {
    if (token == SIZE_MAX) return;
    SVDI_HIP(hipEventRecord(ctx->event_pool[token + 1], ctx->stream));
}

void allow_dynamic_lds(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> granted;
    int device = 0;
    SVDI_HIP(hipGetDevice(&device));
    std::lock_guard<std::mutex> lk(mu);
    size_t& have = granted[{device, kernel}];
    if (bytes <= have) return;
    SVDI_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    have = bytes;
}

// ----------------------------------------------------------------------------- Workspace
void* Workspace::get(Slot s, size_t bytes) {
    if (bytes == 0) bytes = 16;
    bytes = (bytes + 255) & ~static_cast<size_t>(255);
    if (size_[s] >= bytes) return ptr_[s];
    if (frozen_)
        throw ArgError("the workspace would have to grow during stream capture: make one call of this shape on "
                       "the context before capturing");
    if (arena_ != nullptr) {
        if (bytes > arena_cap_ - arena_used_)
            throw ArgError("the call needs more device scratch than the no-allocation entry point's " +
                           std::to_string(arena_cap_ >> 20) + " MiB (" + std::to_string(arena_used_) +
                           " bytes in use, " + std::to_string(bytes) + " more requested)");
        ptr_[s] = arena_ + arena_used_;
        size_[s] = bytes;
        arena_used_ += bytes;
        return ptr_[s];
    }
    if (ptr_[s] != nullptr) {
        SVDI_HIP(hipDeviceSynchronize());
        SVDI_HIP(hipFree(ptr_[s]));
        ptr_[s] = nullptr;
        size_[s] = 0;
    }
    void* p = nullptr;
    SVDI_HIP(hipMalloc(&p, bytes));
    ptr_[s] = p;
    size_[s] = bytes;
    return p;
}

void Workspace::begin_arena(void* base, size_t capacity) {
    for (int i = 0; i < kSlots; ++i) {
        kept_ptr_[i] = ptr_[i];
        kept_size_[i] = size_[i];
        ptr_[i] = nullptr;
        size_[i] = 0;
    }
    arena_ = static_cast<char*>(base);
    arena_cap_ = capacity & ~static_cast<size_t>(255);
    arena_used_ = 0;
}

void Workspace::end_arena() {
    for (int i = 0; i < kSlots; ++i) {
        ptr_[i] = kept_ptr_[i];
        size_[i] = kept_size_[i];
        kept_ptr_[i] = nullptr;
        kept_size_[i] = 0;
    }
    arena_ = nullptr;
    arena_cap_ = arena_used_ = 0;
}

void Workspace::release() {
    for (int i = 0; i < kSlots; ++i) {
        if (ptr_[i] != nullptr) (void)hipFree(ptr_[i]);
        ptr_[i] = nullptr;
        size_[i] = 0;
    }
}

size_t Workspace::bytes_held() const {
    size_t t = 0;
    for (int i = 0; i < kSlots; ++i) t += size_[i];
    return t;
}

// blocking: the context's stream is ordered with the null stream (a caller's context); lanes of
// batched calls take non-blocking streams, ordered only by the fork / join events with their
// parent's stream (so that their work can also be captured into a graph with it).
std::unique_ptr<svdi_context> create_context(int device, bool blocking = true) {
    int count = 0;
    SVDI_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw ArgError("no such HIP device");
    SVDI_HIP(hipSetDevice(device));
    std::unique_ptr<svdi_context> c(new svdi_context());
    c->device = device;
    // A blocking stream: ordered with the null stream both ways, so a caller that works on the
    // null stream (torch's default stream) needs no synchronisation of its own around a call.
    SVDI_HIP(hipStreamCreateWithFlags(&c->own_stream, blocking ? hipStreamDefault : hipStreamNonBlocking));
    c->stream = c->own_stream;
    SVDI_HIP(hipEventCreate(&c->ev_start));
    SVDI_HIP(hipEventCreate(&c->ev_stop));
    SVDI_HIP(hipEventCreateWithFlags(&c->ev_plan, hipEventDisableTiming));
    SVDI_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->pinned), sizeof(PinnedPlan), hipHostMallocDefault));
    std::memset(c->pinned, 0, sizeof(PinnedPlan));
    SVDI_HIP(hipMalloc(reinterpret_cast<void**>(&c->dev_err), 256));
    SVDI_HIP(hipMemsetAsync(c->dev_err, 0, 256, c->own_stream));
    SVDI_HIP(hipStreamSynchronize(c->own_stream));
    hipDeviceProp_t prop;
    SVDI_HIP(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    // A lane's window kernel leaves an eighth of the CUs (four per XCD on MI355X) to the other
    // lanes: 1,306-1,310 against 1,255-1,258 snapshots/s batched (SKA-Low 1024^2; a quarter less
    // or more left free measured lower).
    if (!blocking) c->grid_reserve = c->cu_count / 8;
    c->noalloc_scratch = noalloc_scratch_address(device);
    return c;
}

// A call of the no-allocation entry point: the workspace carves its slots from the device's
// scratch array, the cached grid-correction and twiddle tables are recomputed (the array holds
// whatever the last call on this device left), and no timing events are created.  On leaving,
// the context's own workspace is restored and the stream is drained (also after an error, so
// that no kernel of this call still uses the array).
struct NoAllocCall {
    svdi_context* c;
    bool profiling;
    explicit NoAllocCall(svdi_context* ctx) : c(ctx), profiling(ctx->profiling) {
        c->profiling = false;
        c->conv_at = nullptr;
        c->twiddle_at = nullptr;
        c->ws.begin_arena(c->noalloc_scratch, kNoAllocScratchBytes);
    }
    ~NoAllocCall() {
        (void)hipStreamSynchronize(c->stream);
        c->ws.end_arena();
        c->conv_at = nullptr;
        c->twiddle_at = nullptr;
        c->profiling = profiling;
    }
    NoAllocCall(const NoAllocCall&) = delete;
    NoAllocCall& operator=(const NoAllocCall&) = delete;
};

// The host state of a call on a plan an earlier call reported (svdi_plan_info) for image size N
// and cell `cell`; throws ArgError for a plan no imaging call could have reported.
CallState planned_state(size_t N, float cell, const svdi_plan_info& plan) {
    const size_t G = (N * 3 + 1) / 2;
    if (N == 0 || N > 46340) throw ArgError("Image_Size out of range");
    if (!(cell > 0.0f)) throw ArgError("Cell_Size must be greater than zero");
    if (plan.grid_size != G) throw ArgError("the plan was made for another image size");
    if (plan.num_terms < 1 || plan.num_terms > kMaxTerms || plan.num_slabs < 1 || plan.batch_slabs < 1 ||
        plan.batch_slabs > static_cast<size_t>(plan.num_slabs) || !(plan.slab_width > 0.0f) ||
        !(plan.slab_half_range > 0.0f) || !(plan.vis_det > 0.0f))
        throw ArgError("the plan is not one an imaging call reported");
    CallState cs;
    cs.N = N;
    cs.cell = cell;
    cs.hp.semantics = Semantics::Integrated;
    cs.hp.image_size = N;
    cs.hp.grid_size = G;
    cs.hp.cell = cell;
    DevPlan& d = cs.hp.dev;
    d = DevPlan();
    d.r3_min = plan.r3_min;
    d.r3_max = plan.r3_max;
    d.num_slabs = plan.num_slabs;
    d.slab_width = plan.slab_width;
    d.slab_half_range = plan.slab_half_range;
    d.inv_slab_half_range = plan.inv_slab_half_range;
    d.r3_centre = plan.r3_centre;
    d.slab_alpha_max = plan.slab_alpha_max;
    cs.hp.num_terms = plan.num_terms;
    cs.hp.term_error = plan.term_error;
    cs.hp.batch_slabs = plan.batch_slabs;
    cs.det = plan.vis_det;
    return cs;
}

// For the duration of a planned call: no stage events, and under stream capture a frozen
// workspace (a request for a shape it does not already hold fails instead of allocating).
struct PlannedScope {
    svdi_context* c;
    bool profiling;
    PlannedScope(svdi_context* ctx, bool capturing) : c(ctx), profiling(ctx->profiling) {
        c->profiling = false;
        c->ws.set_frozen(capturing);
    }
    ~PlannedScope() {
        c->ws.set_frozen(false);
        c->profiling = profiling;
    }
    PlannedScope(const PlannedScope&) = delete;
    PlannedScope& operator=(const PlannedScope&) = delete;
};

// Snapshots dealt to L lanes largest first (cost ~ samples + 4 grid cells, the least-loaded lane
// next): queue[j] lists lane j's snapshots in order.
std::vector<std::vector<size_t>> deal_to_lanes(const svdi_snapshot* snaps, size_t count, const size_t* sizes,
                                               size_t L) {
    std::vector<size_t> order(count);
    std::vector<double> cost(count);
    for (size_t i = 0; i < count; ++i) {
        order[i] = i;
        const double G = static_cast<double>((sizes[i] * 3 + 1) / 2);
        cost[i] = static_cast<double>(snaps[i].num_rows) * static_cast<double>(snaps[i].num_channels) + 4.0 * G * G;
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<std::vector<size_t>> queue(L);
    std::vector<double> load(L, 0.0);
    for (size_t i : order) {
        const size_t j = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
        queue[j].push_back(i);
        load[j] += cost[i];
    }
    return queue;
}

// The batched call: enqueue every snapshot on the lanes, then wait for the lanes or not.
// sizes[i] / cells[i]: snapshot i's image side and cell.  Snapshots are dealt to the lanes
// largest first (cost ~ samples + 4 grid cells, the least-loaded lane next), and each lane
// always has its next snapshot's plan queued behind its current work; the host enqueues the
// gridding of whichever lane's plan is back first, waiting only when none is.
// svdi_last_error's message for a failure reported outside this file (the request service).
void report_error(const char* msg) { set_error(msg); }

int image_batch(svdi_context* ctx, const svdi_snapshot* snaps, size_t count, float* const* images,
                const size_t* sizes, const float* cells, const svdi_options* opts, svdi_plan_info* infos,
                bool wait) {
    return guarded([&] {
        if (!ctx || !snaps || !images || !sizes || !cells) throw ArgError("ctx / snapshots / images / sizes / cells is null");
        for (size_t i = 0; i < count; ++i) {
            check_device_pointer(images[i], "image");
            if (sizes[i] == 0) throw ArgError("Image_Size must be greater than zero");
            if (!(cells[i] > 0.0f)) throw ArgError("Cell_Size must be greater than zero");
        }
        if (count == 0) return;
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        const svdi_options o = resolve(opts, SVDI_SEMANTICS_INTEGRATED);
        const size_t L = std::min(count, kBatchLanes);
        while (ctx->lanes.size() < L) ctx->lanes.push_back(create_context(ctx->device, false));
        for (size_t j = 0; j < L; ++j) ctx->lanes[j]->tuning = ctx->tuning;
        const std::vector<std::vector<size_t>> queue = deal_to_lanes(snaps, count, sizes, L);
        // The lanes start after the caller's earlier work on the context stream, and the context
        // stream continues after the last lane's work.
        hipEvent_t ready = nullptr;
        SVDI_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        std::unique_ptr<void, void (*)(void*)> ready_guard(ready, [](void* e) {
            (void)hipEventDestroy(static_cast<hipEvent_t>(e));
        });
        SVDI_HIP(hipEventRecord(ready, ctx->stream));
        for (size_t j = 0; j < L; ++j) SVDI_HIP(hipStreamWaitEvent(ctx->lanes[j]->stream, ready, 0));
        std::vector<CallState> cs(count);
        std::vector<Guess> guesses(count);
        std::vector<SampleInput> in(count);
        std::vector<size_t> next(L, 0);                        // position in the lane's queue
        std::vector<long> planned(L, -1);                      // snapshot whose plan is queued
        size_t order_of_plan = 0;
        std::vector<size_t> plan_seq(L, 0);                    // when that plan was queued
        auto begin_next = [&](size_t j) {
            if (next[j] >= queue[j].size()) return;
            const size_t i = queue[j][next[j]++];
            svdi_context* lane = ctx->lanes[j].get();
            begin_call(lane);
            in[i] = snapshot_input(lane, snaps[i]);
            image_begin(lane, in[i], sizes[i], cells[i], o, cs[i]);
            guesses[i] = guess_and_enqueue(ctx->shapes, lane, in[i], o, cs[i], images[i]);
            planned[j] = static_cast<long>(i);
            plan_seq[j] = order_of_plan++;
        };
        // The lanes' first plans are queued one lane at a time, and between them the host looks
        // for a plan already back, so the first gridding is enqueued as soon as its plan returns
        // rather than after every lane's planning has been enqueued.
        for (;;) {
            long pick = -1;
            for (size_t j = 0; j < L && pick < 0; ++j) {
                if (planned[j] < 0) continue;
                const hipError_t q = hipEventQuery(ctx->lanes[j]->ev_plan);
                if (q == hipSuccess) pick = static_cast<long>(j);
                else if (q != hipErrorNotReady) SVDI_HIP(q);
            }
            if (pick < 0) {
                long idle = -1;                                 // a lane with no plan queued yet
                for (size_t j = 0; j < L && idle < 0; ++j)
                    if (planned[j] < 0 && next[j] < queue[j].size()) idle = static_cast<long>(j);
                if (idle >= 0) {
                    begin_next(static_cast<size_t>(idle));
                    continue;
                }
            }
            if (pick < 0) {                                     // none back yet: the oldest plan
                for (size_t j = 0; j < L; ++j)
                    if (planned[j] >= 0 && (pick < 0 || plan_seq[j] < plan_seq[static_cast<size_t>(pick)]))
                        pick = static_cast<long>(j);
                if (pick < 0) break;                            // every lane is done
                SVDI_HIP(hipEventSynchronize(ctx->lanes[static_cast<size_t>(pick)]->ev_plan));
            }
            const size_t j = static_cast<size_t>(pick);
            const size_t i = static_cast<size_t>(planned[j]);
            image_take_plan(ctx->lanes[j].get(), o, cs[i]);
            settle_guess(ctx->shapes, ctx->lanes[j].get(), in[i], o, guesses[i], cs[i], images[i]);
            planned[j] = -1;
            begin_next(j);
        }
        for (size_t j = 0; j < L; ++j) {
            svdi_context* lane = ctx->lanes[j].get();
            image_finish_enqueue(lane, nullptr, wait);
            SVDI_HIP(hipStreamWaitEvent(ctx->stream, lane->ev_stop, 0));
        }
        // The waiting form reports a binning inconsistency of this call (or of an earlier
        // asynchronous call on the same lanes, whose error words are cleared only when read).
        if (wait)
            for (size_t j = 0; j < L; ++j) {
                SVDI_HIP(hipEventSynchronize(ctx->lanes[j]->ev_stop));
                check_error_word(ctx->lanes[j].get());
            }
        if (infos)
            for (size_t i = 0; i < count; ++i) fill_info(cs[i], 0.0f, &infos[i]);
    });
}

// The batched call on known plans: every snapshot's work is enqueued on the lanes with no wait
// on the host, and the context stream waits for the lanes.  Under stream capture on the context
// stream the lanes' streams join the capture through the fork / join events, so the whole batch
// becomes one graph.
//
// When every snapshot's planes are gridded in one slab batch and the MFMA gridder takes them
// (the common case), the snapshots go in groups of up to kMaxWindowJobs, one lane each: every
// lane bins its snapshot, one window-contraction launch on the context stream takes the whole
// group (persistent workgroups over all the group's windows), and the lanes then transform and
// image their planes.  Otherwise each of up to four lanes takes its snapshots in turn, every
// snapshot gridded on its lane.
int image_batch_planned(svdi_context* ctx, const svdi_snapshot* snaps, size_t count, float* const* images,
                        const size_t* sizes, const float* cells, const svdi_options* opts,
                        const svdi_plan_info* plans) {
    return guarded([&] {
        if (!ctx || !snaps || !images || !sizes || !cells || !plans)
            throw ArgError("ctx / snapshots / images / sizes / cells / plans is null");
        std::vector<CallState> cs(count);
        bool one_pass = true;
        for (size_t i = 0; i < count; ++i) {
            check_device_pointer(images[i], "image");
            cs[i] = planned_state(sizes[i], cells[i], plans[i]);
            one_pass = one_pass && cs[i].hp.batch_slabs >= static_cast<size_t>(cs[i].hp.dev.num_slabs);
        }
        if (count == 0) return;
        const svdi_options o = resolve(opts, SVDI_SEMANTICS_INTEGRATED);
        if (o.semantics != SVDI_SEMANTICS_INTEGRATED) throw ArgError("planned calls take the integrated semantics");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
        SVDI_HIP(hipStreamIsCapturing(ctx->stream, &capture));
        const bool capturing = capture != hipStreamCaptureStatusNone;
        const bool grouped = one_pass && (o.gridder == SVDI_GRIDDER_AUTO || o.gridder == SVDI_GRIDDER_MFMA);
        const size_t L = std::min(count, grouped ? static_cast<size_t>(kMaxWindowJobs) : kBatchLanes);
        if (ctx->lanes.size() < L) {
            if (capturing) throw ArgError("under capture the context needs its lanes from an earlier batched call");
            while (ctx->lanes.size() < L) ctx->lanes.push_back(create_context(ctx->device, false));
        }
        for (size_t j = 0; j < L; ++j) ctx->lanes[j]->tuning = ctx->tuning;
        PlannedScope ctx_scope(ctx, capturing);
        std::vector<std::unique_ptr<PlannedScope>> scopes;
        for (size_t j = 0; j < L; ++j) scopes.emplace_back(new PlannedScope(ctx->lanes[j].get(), capturing));
        // Fork: the lanes start after the caller's earlier work on the context stream (events of
        // the contexts', so that nothing is created under capture).
        SVDI_HIP(hipEventRecord(ctx->ev_plan, ctx->stream));
        for (size_t j = 0; j < L; ++j) SVDI_HIP(hipStreamWaitEvent(ctx->lanes[j]->stream, ctx->ev_plan, 0));
        if (grouped) {
            std::vector<SampleInput> in(L);
            std::vector<Enqueued> e(L);
            for (size_t g0 = 0; g0 < count; g0 += L) {
                const size_t n = std::min(L, count - g0);
                std::vector<WindowJob> jobs;
                std::vector<size_t> binned;                // lanes whose snapshot waits for the launch
                for (size_t k = 0; k < n; ++k) {
                    const size_t i = g0 + k;
                    svdi_context* lane = ctx->lanes[k].get();
                    begin_call(lane);
                    in[k] = snapshot_input(lane, snaps[i]);
                    e[k] = image_prepare(lane, in[k], o, cs[i], images[i]);
                    if (e[k].gridder == SVDI_GRIDDER_MFMA) {
                        e[k].gp.slab_start = 0;
                        e[k].gp.nslabs = static_cast<int>(e[k].total_slabs);
                        jobs.push_back(bin_windows(lane, in[k], e[k].gp, e[k].planes));
                        SVDI_HIP(hipEventRecord(lane->ev_plan, lane->stream));
                        binned.push_back(k);
                    } else {
                        image_grid(lane, in[k], cs[i], e[k], 0, e[k].total_slabs);
                    }
                }
                if (!jobs.empty()) {
                    for (size_t k : binned) SVDI_HIP(hipStreamWaitEvent(ctx->stream, ctx->lanes[k]->ev_plan, 0));
                    contract_windows(ctx, jobs.data(), static_cast<int>(jobs.size()));
                    for (const WindowJob& job : jobs) reduce_windows(ctx, job);
                    SVDI_HIP(hipEventRecord(ctx->ev_plan, ctx->stream));
                    for (size_t k : binned) SVDI_HIP(hipStreamWaitEvent(ctx->lanes[k]->stream, ctx->ev_plan, 0));
                }
                for (size_t k = 0; k < n; ++k)
                    image_post(ctx->lanes[k].get(), cs[g0 + k], e[k], 0, e[k].total_slabs, images[g0 + k]);
            }
        } else {
            const std::vector<std::vector<size_t>> queue = deal_to_lanes(snaps, count, sizes, L);
            for (size_t j = 0; j < L; ++j) {
                svdi_context* lane = ctx->lanes[j].get();
                for (size_t i : queue[j]) {
                    begin_call(lane);
                    SampleInput in = snapshot_input(lane, snaps[i]);
                    image_enqueue(lane, in, o, cs[i], images[i]);
                }
            }
        }
        // Join: the context stream continues after every lane's last snapshot.
        for (size_t j = 0; j < L; ++j) {
            svdi_context* lane = ctx->lanes[j].get();
            SVDI_HIP(hipEventRecord(lane->ev_plan, lane->stream));
            SVDI_HIP(hipStreamWaitEvent(ctx->stream, lane->ev_plan, 0));
        }
    });
}

}  // namespace svdi

svdi_context::~svdi_context() {
    (void)hipSetDevice(device);
    lanes.clear();
    if (stream) (void)hipStreamSynchronize(stream);
    if (own_stream && own_stream != stream) (void)hipStreamSynchronize(own_stream);
    ws.release();
    if (pinned) (void)hipHostFree(pinned);
    if (dev_err) (void)hipFree(dev_err);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (ev_plan) (void)hipEventDestroy(ev_plan);
    if (own_stream) (void)hipStreamDestroy(own_stream);
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
}

using namespace svdi;

extern "C" {

int svdi_version(void) { return SVDI_ABI_VERSION; }

const char* svdi_last_error(void) { return g_last_error.c_str(); }

void svdi_default_options(svdi_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->gridder = SVDI_GRIDDER_AUTO;
    o->output = SVDI_OUTPUT_DIRTY;
    o->semantics = SVDI_SEMANTICS_INTEGRATED;
}

int svdi_create(svdi_context** out, int device) {
    return guarded([&] {
        if (!out) throw ArgError("ctx is null");
        *out = nullptr;
        *out = create_context(device).release();
    });
}

int svdi_destroy(svdi_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        delete ctx;
    });
}

int svdi_set_stream(svdi_context* ctx, void* stream) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    });
}

int svdi_trim(svdi_context* ctx) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        SVDI_HIP(hipStreamSynchronize(ctx->stream));
        for (auto& lane : ctx->lanes) {
            SVDI_HIP(hipStreamSynchronize(lane->stream));
            lane->ws.release();
            lane->twiddle_n = 0;
            lane->conv_at = nullptr;
        }
        ctx->ws.release();
        ctx->twiddle_n = 0;
        ctx->conv_at = nullptr;
    });
}

int svdi_workspace_bytes(svdi_context* ctx, size_t* bytes) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        if (!bytes) throw ArgError("bytes is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        size_t t = ctx->ws.bytes_held();
        for (const auto& lane : ctx->lanes) t += lane->ws.bytes_held();
        *bytes = t;
    });
}

int svdi_preprocess(svdi_context* ctx, const svdi_snapshot* snap, float* bin, float* vin,
                    float* vis_re, float* vis_im) {
    return guarded([&] {
        if (!ctx || !snap) throw ArgError("ctx / snapshot is null");
        check_device_pointer(bin, "bin");
        check_device_pointer(vin, "vin");
        check_device_pointer(vis_re, "vis_re");
        check_device_pointer(vis_im, "vis_im");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        SampleInput in = snapshot_input(ctx, *snap);
        launch_materialize(ctx, *snap, in.geo, bin, vin, vis_re, vis_im);
        SVDI_HIP(hipStreamSynchronize(ctx->stream));
        finish_call(ctx);
    });
}

int svdi_snapshot_geometry(svdi_context* ctx, const svdi_snapshot* snap, double* rot, double* evals) {
    return guarded([&] {
        if (!ctx || !snap || !rot) throw ArgError("ctx / snapshot / rot is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        SampleInput in = snapshot_input(ctx, *snap);
        DevGeometry host;
        SVDI_HIP(hipMemcpyAsync(&host, in.geo, sizeof host, hipMemcpyDeviceToHost, ctx->stream));
        SVDI_HIP(hipStreamSynchronize(ctx->stream));
        finish_call(ctx);
        std::memcpy(rot, host.R, sizeof host.R);
        if (evals) std::memcpy(evals, host.evals, sizeof host.evals);
    });
}

int svdi_image_preprocessed(svdi_context* ctx, const float* bin, const float* vin,
                            const float* vis_re, const float* vis_im, size_t num_samples,
                            float* image, size_t image_size, float cell_size,
                            const svdi_options* opts, svdi_plan_info* info) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        check_device_pointer(bin, "bin");
        check_device_pointer(vin, "vin");
        check_device_pointer(vis_re, "vis_re");
        check_device_pointer(vis_im, "vis_im");
        check_device_pointer(image, "image");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        SampleInput in;
        in.num_samples = num_samples;
        in.bin = bin;
        in.vis_re = vis_re;
        in.vis_im = vis_im;
        in.vin_dev = vin;
        image_core(ctx, in, image_size, cell_size, resolve(opts, SVDI_SEMANTICS_INTEGRATED), info, image);
    });
}

int svdi_image_snapshot(svdi_context* ctx, const svdi_snapshot* snap, float* image,
                        size_t image_size, float cell_size, const svdi_options* opts,
                        svdi_plan_info* info) {
    return guarded([&] {
        if (!ctx || !snap) throw ArgError("ctx / snapshot is null");
        check_device_pointer(image, "image");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        SampleInput in = snapshot_input(ctx, *snap);
        image_core(ctx, in, image_size, cell_size, resolve(opts, SVDI_SEMANTICS_INTEGRATED), info, image, true);
    });
}


int svdi_image_snapshot_noalloc(svdi_context* ctx, const svdi_snapshot* snap, float* image,
                                size_t image_size, float cell_size, const svdi_options* opts,
                                svdi_plan_info* info) {
    return guarded([&] {
        if (!ctx || !snap) throw ArgError("ctx / snapshot is null");
        check_device_pointer(image, "image");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        if (ctx->noalloc_scratch == nullptr) throw std::runtime_error("the context has no scratch address");
        begin_call(ctx);
        NoAllocCall scope(ctx);
        SampleInput in = snapshot_input(ctx, *snap);
        svdi_options o = resolve(opts, SVDI_SEMANTICS_INTEGRATED);
        if (o.max_batch_slabs == 0) o.max_batch_slabs = 1;   // one slab's planes at a time
        image_core(ctx, in, image_size, cell_size, o, info, image);
    });
}

int svdi_image_snapshot_planned(svdi_context* ctx, const svdi_snapshot* snap, float* image,
                                size_t image_size, float cell_size, const svdi_options* opts,
                                const svdi_plan_info* plan) {
    return guarded([&] {
        if (!ctx || !snap || !plan) throw ArgError("ctx / snapshot / plan is null");
        check_device_pointer(image, "image");
        CallState cs = planned_state(image_size, cell_size, *plan);
        svdi_options o = resolve(opts, SVDI_SEMANTICS_INTEGRATED);
        if (o.semantics != SVDI_SEMANTICS_INTEGRATED) throw ArgError("planned calls take the integrated semantics");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        // Under stream capture the workspace must already hold this shape (no allocation).
        hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
        SVDI_HIP(hipStreamIsCapturing(ctx->stream, &capture));
        PlannedScope scope(ctx, capture != hipStreamCaptureStatusNone);
        begin_call(ctx);
        SampleInput in = snapshot_input(ctx, *snap);
        image_enqueue(ctx, in, o, cs, image);
    });
}

int svdi_image_snapshots_planned(svdi_context* ctx, const svdi_snapshot* snapshots, size_t count,
                                 float* const* images, const size_t* image_sizes, const float* cell_sizes,
                                 const svdi_options* opts, const svdi_plan_info* plans) {
    return svdi::image_batch_planned(ctx, snapshots, count, images, image_sizes, cell_sizes, opts, plans);
}

int svdi_image_snapshot_batch(svdi_context* ctx, const svdi_snapshot* snaps, size_t count,
                              float* const* images, size_t image_size, float cell_size,
                              const svdi_options* opts, svdi_plan_info* infos) {
    std::vector<size_t> sizes;
    std::vector<float> cells;
    try {
        sizes.assign(count, image_size);
        cells.assign(count, cell_size);
    } catch (...) {
        set_error("host allocation failed");
        return SVDI_ERR_OUT_OF_MEMORY;
    }
    return svdi::image_batch(ctx, snaps, count, images, sizes.data(), cells.data(), opts, infos, true);
}

int svdi_image_snapshot_batch_async(svdi_context* ctx, const svdi_snapshot* snaps, size_t count,
                                    float* const* images, size_t image_size, float cell_size,
                                    const svdi_options* opts, svdi_plan_info* infos) {
    std::vector<size_t> sizes;
    std::vector<float> cells;
    try {
        sizes.assign(count, image_size);
        cells.assign(count, cell_size);
    } catch (...) {
        set_error("host allocation failed");
        return SVDI_ERR_OUT_OF_MEMORY;
    }
    return svdi::image_batch(ctx, snaps, count, images, sizes.data(), cells.data(), opts, infos, false);
}

int svdi_image_snapshots(svdi_context* ctx, const svdi_snapshot* snaps, size_t count, float* const* images,
                         const size_t* image_sizes, const float* cell_sizes, const svdi_options* opts,
                         svdi_plan_info* infos, int wait) {
    return svdi::image_batch(ctx, snaps, count, images, image_sizes, cell_sizes, opts, infos, wait != 0);
}

int svdi_fipipe_host(svdi_context* ctx, const float* vis_re, const float* vis_im,
                     const float* bin, const float* vin, float* dirty_image,
                     size_t num_baselines, size_t image_size, float cell_size,
                     const svdi_options* opts, svdi_plan_info* info) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        if (!vis_re || !vis_im || !bin || !vin || !dirty_image) throw ArgError("null host buffer");
        if (num_baselines == 0) throw ArgError("num_baselines must be greater than zero");
        if (image_size == 0) throw ArgError("Image_Size must be greater than zero");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        const size_t S = num_baselines;
        float* d_bin = static_cast<float*>(ctx->ws.get(Workspace::kInBin, S * 3 * sizeof(float)));
        float* d_vre = static_cast<float*>(ctx->ws.get(Workspace::kInVre, S * sizeof(float)));
        float* d_vim = static_cast<float*>(ctx->ws.get(Workspace::kInVim, S * sizeof(float)));
        float* d_vin = static_cast<float*>(ctx->ws.get(Workspace::kInVin, 9 * sizeof(float)));
        float* d_img = static_cast<float*>(ctx->ws.get(Workspace::kImage, image_size * image_size * sizeof(float)));
        SVDI_HIP(hipMemcpyAsync(d_bin, bin, S * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        SVDI_HIP(hipMemcpyAsync(d_vre, vis_re, S * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        SVDI_HIP(hipMemcpyAsync(d_vim, vis_im, S * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        SVDI_HIP(hipMemcpyAsync(d_vin, vin, 9 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        SampleInput in;
        in.num_samples = S;
        in.bin = d_bin;
        in.vis_re = d_vre;
        in.vis_im = d_vim;
        in.vin_dev = d_vin;
        image_core(ctx, in, image_size, cell_size, resolve(opts, SVDI_SEMANTICS_ARCHIVE), info, d_img);
        SVDI_HIP(hipMemcpyAsync(dirty_image, d_img, image_size * image_size * sizeof(float),
                                hipMemcpyDeviceToHost, ctx->stream));
        SVDI_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int svdi_ms_read_snapshot(const char* column_dir, size_t snapshot_index, svdi_ms_snapshot* out) {
    if (!column_dir || !out) {
        svdi::set_error("column_dir / out is null");
        return SVDI_ERR_INVALID_ARGUMENT;
    }
    *out = svdi_ms_snapshot{};
    std::unique_ptr<svdi::MsSnapshot> s;
    try {
        s.reset(new svdi::MsSnapshot(svdi::read_ms_snapshot(column_dir, snapshot_index)));
    } catch (const std::bad_alloc&) {
        svdi::set_error("host allocation failed");
        return SVDI_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        svdi::set_error(e.what());
        return SVDI_ERR_IO;
    }
    svdi_snapshot& d = out->data;
    d.num_rows = s->num_rows;
    d.num_channels = s->num_channels;
    d.uvw = s->uvw.data();
    d.freq_hz = s->freq_hz.data();
    d.vis0_re = s->vis0_re.data();
    d.vis0_im = s->vis0_im.data();
    d.vis3_re = s->vis3_re.data();
    d.vis3_im = s->vis3_im.data();
    d.flag0 = s->flag0.data();
    d.flag3 = s->flag3.data();
    d.weight0 = s->has_weights ? s->weight0.data() : nullptr;
    d.weight3 = s->has_weights ? s->weight3.data() : nullptr;
    d.array_center_x = s->array_center_x;
    out->index = s->index;
    out->total_snapshots = s->total_snapshots;
    out->time = s->time;
    out->internal = s.release();
    return SVDI_OK;
}

void svdi_ms_free_snapshot(svdi_ms_snapshot* snapshot) {
    if (!snapshot) return;
    delete static_cast<svdi::MsSnapshot*>(snapshot->internal);
    *snapshot = svdi_ms_snapshot{};
}

int svdi_image_host_snapshot(svdi_context* ctx, const svdi_snapshot* hs, float* image, size_t image_size,
                             float cell_size, const svdi_options* opts, svdi_plan_info* info) {
    return guarded([&] {
        if (!ctx || !hs || !image) throw ArgError("ctx / snapshot / image is null");
        if (hs->num_rows == 0 || hs->num_channels == 0) throw ArgError("Measurement set is empty");
        if (!hs->uvw || !hs->freq_hz || !hs->vis0_re || !hs->vis0_im || !hs->vis3_re || !hs->vis3_im ||
            !hs->flag0 || !hs->flag3)
            throw ArgError("null host array in the snapshot");
        if (image_size == 0) throw ArgError("Image_Size must be greater than zero");
        if (image_size > 46340) throw ArgError("Image_Size too large");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        const size_t R = hs->num_rows, C = hs->num_channels, S = R * C;
        // The arrays in one workspace slot, each at a 256-byte boundary (a NULL weight array
        // stays NULL: weights 1, as on the device path).
        auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
        const size_t sizes[10] = {3 * R * sizeof(double), C * sizeof(double), S * sizeof(float), S * sizeof(float),
                                  S * sizeof(float), S * sizeof(float), S, S,
                                  hs->weight0 ? S * sizeof(float) : 0, hs->weight3 ? S * sizeof(float) : 0};
        const void* srcs[10] = {hs->uvw, hs->freq_hz, hs->vis0_re, hs->vis0_im, hs->vis3_re, hs->vis3_im,
                                hs->flag0, hs->flag3, hs->weight0, hs->weight3};
        size_t total = 0;
        for (int i = 0; i < 10; ++i) total += up(sizes[i]);
        char* base = static_cast<char*>(ctx->ws.get(Workspace::kInSnap, total));
        void* dst[10];
        size_t off = 0;
        for (int i = 0; i < 10; ++i) {
            dst[i] = sizes[i] ? base + off : nullptr;
            if (sizes[i]) SVDI_HIP(hipMemcpyAsync(dst[i], srcs[i], sizes[i], hipMemcpyHostToDevice, ctx->stream));
            off += up(sizes[i]);
        }
        float* d_img = static_cast<float*>(ctx->ws.get(Workspace::kImage, image_size * image_size * sizeof(float)));
        svdi_snapshot ds = *hs;
        ds.uvw = static_cast<const double*>(dst[0]);
        ds.freq_hz = static_cast<const double*>(dst[1]);
        ds.vis0_re = static_cast<const float*>(dst[2]);
        ds.vis0_im = static_cast<const float*>(dst[3]);
        ds.vis3_re = static_cast<const float*>(dst[4]);
        ds.vis3_im = static_cast<const float*>(dst[5]);
        ds.flag0 = static_cast<const uint8_t*>(dst[6]);
        ds.flag3 = static_cast<const uint8_t*>(dst[7]);
        ds.weight0 = static_cast<const float*>(dst[8]);
        ds.weight3 = static_cast<const float*>(dst[9]);
        SampleInput in = snapshot_input(ctx, ds);
        image_core(ctx, in, image_size, cell_size, resolve(opts, SVDI_SEMANTICS_INTEGRATED), info, d_img);
        SVDI_HIP(hipMemcpyAsync(image, d_img, image_size * image_size * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
        SVDI_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int svdi_fft2_inverse(svdi_context* ctx, const void* planes, size_t num_planes, size_t grid_size,
                      size_t row_lo, size_t rows, size_t col_lo, size_t cols, void* out) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        check_device_pointer(planes, "planes");
        check_device_pointer(out, "out");
        if (num_planes == 0 || grid_size == 0 || rows == 0 || cols == 0) throw ArgError("empty transform");
        if (grid_size > 69510) throw ArgError("grid_size too large");
        if (row_lo + rows > grid_size || col_lo + cols > grid_size) throw ArgError("window outside the plane");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        begin_call(ctx);
        float2* rowfft = static_cast<float2*>(
            ctx->ws.get(Workspace::kRowFft, num_planes * grid_size * cols * sizeof(float2)));
        launch_fft2_inverse(ctx, static_cast<const float2*>(planes), num_planes, grid_size, row_lo, rows, col_lo,
                            cols, rowfft, static_cast<float2*>(out));
        SVDI_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int svdi_set_tuning(svdi_context* ctx, size_t grid_size, const svdi_tuning* tuning) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        if (tuning) {
            if (tuning->coop_min != 0 && (tuning->coop_min < 32 || tuning->coop_min > 65536))
                throw ArgError("coop_min must be 0 or in 32 .. 65536");
            if (tuning->coop_cap != 0 && (tuning->coop_cap < 4096 || tuning->coop_cap > (1u << 24)))
                throw ArgError("coop_cap must be 0 or in 4096 .. 16777216");
            if (tuning->class_log2 != 0 && tuning->class_log2 != 2 && tuning->class_log2 != 3)
                throw ArgError("class_log2 must be 0, 2 or 3");
        }
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (tuning) ctx->tuning[grid_size] = *tuning;
        else ctx->tuning.erase(grid_size);
    });
}

int svdi_get_tuning(svdi_context* ctx, size_t grid_size, svdi_tuning* tuning) {
    return guarded([&] {
        if (!ctx || !tuning) throw ArgError("ctx / tuning is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        *tuning = tuning_for(ctx, grid_size);
    });
}

int svdi_autotune(svdi_context* ctx, const svdi_snapshot* snap, size_t image_size, float cell_size,
                  const svdi_options* opts, int repeats, svdi_tuning* best) {
    return guarded([&] {
        if (!ctx || !snap || !best) throw ArgError("ctx / snapshot / best is null");
        if (repeats < 1 || repeats > 100) throw ArgError("repeats must be in 1 .. 100");
        const size_t N = image_size, G = (N * 3 + 1) / 2;
        if (N == 0 || N > 46340) throw ArgError("Image_Size out of range");
        if (!(cell_size > 0.0f)) throw ArgError("Cell_Size must be greater than zero");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        svdi_options o = resolve(opts, SVDI_SEMANTICS_INTEGRATED);
        o.gridder = SVDI_GRIDDER_MFMA;
        o.verbose = 0;
        float* image = nullptr;
        SVDI_HIP(hipMalloc(&image, N * N * sizeof(float)));
        struct Restore {
            svdi_context* c;
            float* image;
            bool profiling;
            ~Restore() {
                (void)hipStreamSynchronize(c->stream);
                (void)hipFree(image);
                c->profiling = profiling;
            }
        } restore{ctx, image, ctx->profiling};
        ctx->profiling = true;
        const bool had_entry = ctx->tuning.count(G) != 0;
        const svdi_tuning saved = had_entry ? ctx->tuning[G] : svdi_tuning{};
        // The grid stage's best time over `repeats` calls with shape t (stage [2]: binning, item
        // lists, the window contraction and the piece sums).
        auto time_shape = [&](const svdi_tuning& t) {
            ctx->tuning[G] = t;
            float ms = 0.0f;
            for (int r = 0; r < repeats; ++r) {
                begin_call(ctx);
                SampleInput in = snapshot_input(ctx, *snap);
                image_core(ctx, in, N, cell_size, o, nullptr, image);
                const float m = ctx->stage_ms[2];
                if (r == 0 || m < ms) ms = m;
            }
            return ms;
        };
        try {
            const svdi_tuning defaults = {};
            (void)time_shape(defaults);                 // the workspace grows to this shape first
            const float default_ms = time_shape(defaults);
            svdi_tuning pick = defaults;
            pick.grid_ms = default_ms;
            for (uint32_t cmin : {256u, 512u, 1024u, 2048u})
                for (uint32_t ccap : {12288u, 24576u, 49152u})
                    for (int cls : {2, 3}) {
                        svdi_tuning t = {};
                        t.coop_min = cmin;
                        t.coop_cap = ccap;
                        t.class_log2 = cls;
                        const float ms = time_shape(t);
                        if (ms < pick.grid_ms) {
                            pick = t;
                            pick.grid_ms = ms;
                        }
                    }
            // The defaults once more (the first timings of a context run cold): a shape replaces
            // them only when it beats their better time.
            const float default_ms2 = std::min(default_ms, time_shape(defaults));
            if (!(pick.grid_ms < default_ms2)) {
                pick = defaults;
                pick.grid_ms = default_ms2;
            }
            pick.default_ms = default_ms2;
            ctx->tuning[G] = pick;
            *best = pick;
        } catch (...) {
            if (had_entry) ctx->tuning[G] = saved;
            else ctx->tuning.erase(G);
            throw;
        }
    });
}

int svdi_set_profiling(svdi_context* ctx, int enable) {
    return guarded([&] {
        if (!ctx) throw ArgError("ctx is null");
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->profiling = enable != 0;
    });
}

int svdi_last_stage_times(svdi_context* ctx, float* ms, int count) {
    return guarded([&] {
        if (!ctx || !ms) throw ArgError("null argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (int i = 0; i < count; ++i) ms[i] = i < SVDI_NUM_STAGES ? ctx->stage_ms[i] : 0.0f;
    });
}

void svdi_free(void* p) { std::free(p); }

}  // extern "C"

// archive/pipe_interface.cpp:8 — the C++-linkage entry point the archive CLI links against.
// One process-wide context, created on first use, serialises calls.
int FIpipe(float* Visreal, float* Visimag, float* Bin, float* Vin, float* dirty_image,
           size_t num_baselines, size_t image_size, float cell_size) {
    static std::mutex mu;
    static svdi_context* ctx = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (ctx == nullptr) {
        int dev = 0;
        if (const char* e = std::getenv("SVDIMAGER_DEVICE")) dev = std::atoi(e);
        if (svdi_create(&ctx, dev) != SVDI_OK) {
            std::fprintf(stderr, "FIpipe: %s\n", svdi_last_error());
            ctx = nullptr;
            return 1;
        }
    }
    svdi_options o;
    svdi_default_options(&o);
    o.semantics = SVDI_SEMANTICS_ARCHIVE;
    o.verbose = 1;
    const int rc = svdi_fipipe_host(ctx, Visreal, Visimag, Bin, Vin, dirty_image, num_baselines,
                                    image_size, cell_size, &o, nullptr);
    if (rc != SVDI_OK) {
        std::fprintf(stderr, "Error sync : %s\n", svdi_last_error());
        return rc;
    }
    return 0;
}
