/* libsvdimager — MI355X (gfx950) snapshot imager behind the egbdfX/SVDimager host API.
 *
 * Pipeline: SVD rotation of uvw -> flag / WEIGHT_SPECTRUM polarisation collapse with MFS channel
 * folding -> exponential-of-semicircle gridding of Chebyshev w-moments -> batched 2D inverse
 * FFT -> image-plane Chebyshev w-phase recombination with ES grid correction -> SIN
 * reprojection splat -> dirty image.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never throws, and returns
 * an svdi_status.  svdi_last_error() gives the message of the last failure on the calling
 * thread.  Device pointers are HIP device memory of the context's device.
 *
 * The C++-linkage symbol FIpipe (archive/pipe_interface.cpp:8) is exported too, with the
 * exact signature the reference's archive CLI links against; it is declared at the end of
 * this header for C++ callers.
 */
#ifndef SVDIMAGER_H
#define SVDIMAGER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVDI_ABI_VERSION 1

typedef enum svdi_status {
    SVDI_OK = 0,
    SVDI_ERR_INVALID_ARGUMENT = 1,
    SVDI_ERR_HIP = 2,
    SVDI_ERR_OUT_OF_MEMORY = 3,
    SVDI_ERR_UNSUPPORTED = 4,
    SVDI_ERR_IO = 5,
    SVDI_ERR_INTERNAL = 6
} svdi_status;

typedef struct svdi_context svdi_context;

/* One TIME snapshot, the fields of HostMeasurementSetData (SVDIntegratedImager.cu:39-54) with
 * uvw and channel frequencies kept in FP64.  Sample index = row + chan * num_rows
 * (SVDIntegratedImager.cu:90-92).  weight0 / weight3 may be NULL when WEIGHT_SPECTRUM is
 * absent (weights 1, SVDIntegratedImager.cu:252-253).  All pointers are device pointers. */
typedef struct svdi_snapshot {
    size_t num_rows;
    size_t num_channels;
    const double* uvw;          /* num_rows x 3, metres */
    const double* freq_hz;      /* num_channels */
    const float* vis0_re;
    const float* vis0_im;
    const float* vis3_re;
    const float* vis3_im;
    const uint8_t* flag0;
    const uint8_t* flag3;
    const float* weight0;
    const float* weight3;
    double array_center_x;      /* OBSERVATION/ARRAY_CENTER x, metres (0 = unknown) */
} svdi_snapshot;

typedef enum svdi_gridder {
    SVDI_GRIDDER_AUTO = 0,      /* MFMA, or TILED for grids and sample counts MFMA does not take */
    SVDI_GRIDDER_ATOMIC = 1,    /* one thread per sample, device atomics (the reference's scheme) */
    SVDI_GRIDDER_TILED = 2,     /* samples binned into LDS-resident grid tiles, FP32 */
    SVDI_GRIDDER_MFMA = 3       /* 16 x 16 output windows held in MFMA accumulators, f16 hi/lo
                                   operands, FP32 accumulation */
} svdi_gridder;

typedef enum svdi_output {
    SVDI_OUTPUT_DIRTY = 0,      /* |.| of each slab batch, bilinearly splatted (the reference's image) */
    SVDI_OUTPUT_SIGNED = 1      /* signed, grid-corrected image before |.| and reprojection, x 1/S */
} svdi_output;

typedef enum svdi_semantics {
    SVDI_SEMANTICS_INTEGRATED = 0,  /* SVDIntegratedImager.cu: slabs, K in {8,16,24,32} */
    SVDI_SEMANTICS_ARCHIVE = 1      /* archive/FIPkernels.cu: one slab, smallest K in 1..32 */
} svdi_semantics;

/* Operand precision of the MFMA gridder's window contraction D += C B (C: the visibility times
 * the Chebyshev term, B: the product of the two ES factors), accumulated in FP32. */
typedef enum svdi_mfma_precision {
    SVDI_PRECISION_DEFAULT = 0,     /* the library's choice: SPLIT3 */
    SVDI_PRECISION_SPLIT3 = 1,      /* C and both ES factors in f16 hi + lo, B's rounding kept:
                                       3 MFMAs per window row, ~22 significant bits per product */
    SVDI_PRECISION_SPLITC = 2,      /* C in f16 hi + lo, B a single f16: 2 MFMAs per row */
    SVDI_PRECISION_SINGLE = 3,      /* C and B single f16: 1 MFMA per row */
    SVDI_PRECISION_BF16 = 4         /* C and B in bf16 hi + lo (B's product formed in FP32):
                                       3 bf16 MFMAs per row, ~16 significant bits per product */
} svdi_mfma_precision;

typedef struct svdi_options {
    int gridder;            /* svdi_gridder */
    int output;             /* svdi_output */
    int semantics;          /* svdi_semantics */
    int num_terms;          /* 0 = the reference's selection, else a fixed K in 1..32 */
    float target_alpha;     /* <= 0: SVD_CHEB_TARGET_ALPHA from the environment, else 10 */
    size_t max_batch_slabs; /* 0: SVDIMAGER_MAX_BATCH_SLABS from the environment, else 25% of free HBM */
    int verbose;            /* 1: print the reference's Chebyshev report and "Time elapsed" to stdout */
    int mfma_precision;     /* an svdi_mfma_precision value; the MFMA gridder only */
} svdi_options;

/* Planning results of one imaging call. */
typedef struct svdi_plan_info {
    float r3_min, r3_max, max_abs_t3;
    uint32_t valid_pixels;
    float alpha_max;        /* Phi_max */
    float target_alpha;
    int num_slabs;
    float slab_width, slab_half_range, slab_alpha_max;
    int num_terms;          /* selected K */
    float term_error;       /* sampled Chebyshev error of the selected K */
    float bucket_errors[4]; /* integrated semantics: errors at K = 8, 16, 24, 32 */
    size_t batch_slabs;
    size_t grid_size;
    size_t num_planes;      /* slabs x K */
    size_t records;         /* records gridded by the tiled or MFMA gridder (last slab batch) */
    float elapsed_ms;       /* device time from planning to the finished image */
    int gridder;            /* svdi_gridder that ran (AUTO resolved) */
    float r3_centre;        /* (r3_min + r3_max) / 2 */
    float inv_slab_half_range;
    float vis_det;          /* |V00 V11 - V01 V10| of the rotation rows (the visibility scale) */
    int enqueued_early;     /* svdi_image_snapshot: 1 = enqueued whole before the plan was read back,
                               on the K of the context's last one-slab call of this size, cell and
                               options, and kept; 2 = that guess missed and the work was enqueued
                               again with the plan's shape; 0 = enqueued after the read-back */
} svdi_plan_info;

int svdi_version(void);
const char* svdi_last_error(void);
void svdi_default_options(svdi_options* opts);

/* Context: one HIP device, one stream, a workspace reused across calls.  Calls on one context
 * are serialised; use one context per host thread for concurrency.  The context's own stream
 * is a blocking stream, ordered with the null stream in both directions: work the caller put on
 * the null stream before a call is complete before the call reads its inputs, and null-stream
 * work after an asynchronous call sees its outputs.  Batched calls run on lanes that fork from
 * the context stream and that it joins (events), so the same holds for them. */
int svdi_create(svdi_context** ctx, int device);
int svdi_destroy(svdi_context* ctx);
/* Run subsequent work on a caller stream (hipStream_t); NULL restores the context's stream. */
int svdi_set_stream(svdi_context* ctx, void* hip_stream);
/* Release the workspace, the batch lanes' included (it regrows on the next call). */
int svdi_trim(svdi_context* ctx);
/* Device bytes the context's workspace holds, the batch lanes' included (no counterpart in the
 * reference, which allocates and frees inside every call, SVDIntegratedImager.cu:1498-1761).
 * The workspace only grows, so after calls at the largest shapes a caller knows what the
 * context keeps and can budget the rest of device memory. */
int svdi_workspace_bytes(svdi_context* ctx, size_t* bytes);

/* Stage timing with HIP events on the context stream (no counterpart in the reference, which
 * times only the whole slab loop, SVDIntegratedImager.cu:1657-1741).  When enabled, each call
 * records events around its stages; svdi_last_stage_times returns the device milliseconds of
 * the last call per stage: [0] preprocess, [1] plan, [2] grid, [3] fft, [4] image, and
 * [5] the accumulation kernel alone, a part of [2] (the binning and sort are the rest). */
#define SVDI_NUM_STAGES 6
int svdi_set_profiling(svdi_context* ctx, int enable);
int svdi_last_stage_times(svdi_context* ctx, float* ms, int count);

/* preprocess_measurement_set_gpu (SVDIntegratedImager.cu:635-853): device snapshot ->
 * Bin (S x 3, rotated, uncentred, wavelengths), Vin (3 x 3 row-major), collapsed vis (S each).
 * Output buffers are caller-allocated device memory.  The rotation is computed in FP64. */
int svdi_preprocess(svdi_context* ctx, const svdi_snapshot* snapshot, float* bin, float* vin,
                    float* vis_re, float* vis_im);

/* The FP64 rotation the snapshot imager uses (rows = principal axes after the orientation rules
 * of SVDIntegratedImager.cu:538-586; the reference keeps it only in FP32 as Vin) and the
 * ascending eigenvalues of the baseline scatter.  rot: host double[9], row-major; evals: host
 * double[3] or NULL. */
int svdi_snapshot_geometry(svdi_context* ctx, const svdi_snapshot* snapshot, double* rot,
                           double* evals);

/* FIpipeDevice (SVDIntegratedImager.cu:1498-1761): preprocessed device buffers -> image_size^2
 * float image in device memory (row-major, x fastest).  opts / info may be NULL. */
int svdi_image_preprocessed(svdi_context* ctx, const float* bin, const float* vin,
                            const float* vis_re, const float* vis_im, size_t num_samples,
                            float* image, size_t image_size, float cell_size,
                            const svdi_options* opts, svdi_plan_info* info);

/* preprocess_measurement_set_gpu followed by FIpipeDevice (SVDIntegratedImager.cu:1822-1828)
 * without materialising Bin: positions are rotated in FP64 and visibilities collapsed inside
 * the gridding pass. */
int svdi_image_snapshot(svdi_context* ctx, const svdi_snapshot* snapshot, float* image,
                        size_t image_size, float cell_size, const svdi_options* opts,
                        svdi_plan_info* info);

/* svdi_image_snapshot without any device allocation (no counterpart in the reference; for
 * callers that cannot allocate: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The call's scratch (rotation, plan, binning records, planes, FFT buffers)
 * is carved from a fixed device array of SVDI_NOALLOC_SCRATCH_BYTES that the library holds on
 * every GPU it runs on until the process ends; a snapshot or image that needs more returns
 * SVDI_ERR_INVALID_ARGUMENT with the sizes in svdi_last_error() (about 2 M samples at 512^2
 * fit).  Slabs are imaged one at a time unless opts->max_batch_slabs says otherwise.  Every
 * context on a GPU shares that GPU's array: calls that overlap in time on one GPU (from two
 * contexts or threads) race and must be serialised by the caller.  The call returns when the
 * image is complete. */
#define SVDI_NOALLOC_SCRATCH_BYTES (256u << 20)
int svdi_image_snapshot_noalloc(svdi_context* ctx, const svdi_snapshot* snapshot, float* image,
                                size_t image_size, float cell_size, const svdi_options* opts,
                                svdi_plan_info* info);

/* svdi_image_snapshot on a plan an earlier call reported for the same geometry (no counterpart in
 * the reference; for re-imaging fixed uvw and frequencies with new visibilities, and for HIP
 * graph capture): the planning kernels and the host's wait for their result are skipped, the
 * plan's slabs, K and slab batches are used as given, and the call returns once the work is
 * enqueued on the context's stream (the image is complete when the stream reaches that point).
 * Nothing in the call waits on the host or records a memset, so it can be captured into a
 * graph on the context's stream (svdi_set_stream): under capture the workspace must already
 * hold the shape (one earlier call of the same sizes on the context), else the call fails with
 * SVDI_ERR_INVALID_ARGUMENT instead of allocating.  The plan must come from a call on the same
 * uvw, frequencies and flags, image size and cell: samples outside its r3 range would be
 * imaged with the wrong w-terms. */
int svdi_image_snapshot_planned(svdi_context* ctx, const svdi_snapshot* snapshot, float* image,
                                size_t image_size, float cell_size, const svdi_options* opts,
                                const svdi_plan_info* plan);

/* Batched svdi_image_snapshot (no counterpart in the reference, which images one snapshot per
 * process): `count` device-resident snapshots imaged into images[i] (device pointers,
 * image_size^2 floats each), up to four at a time on concurrent streams with their own
 * workspaces, so that one snapshot's latency-bound phases overlap another's.  Work starts after
 * the caller's earlier work on the context stream and the call returns when every image is
 * done.  infos (count entries) may be NULL; their elapsed_ms and records are 0. */
int svdi_image_snapshot_batch(svdi_context* ctx, const svdi_snapshot* snapshots, size_t count,
                              float* const* images, size_t image_size, float cell_size,
                              const svdi_options* opts, svdi_plan_info* infos);

/* svdi_image_snapshot_batch without the final wait (no counterpart in the reference; for callers
 * that pipeline batches): the call returns once the last snapshot's work is enqueued, which is
 * after every snapshot's plan is known.  The context stream (the caller's stream when one was
 * set with svdi_set_stream) waits for all of the call's work, so the images are complete once
 * that stream reaches this point; snapshots and images must stay allocated until then. */
int svdi_image_snapshot_batch_async(svdi_context* ctx, const svdi_snapshot* snapshots, size_t count,
                                    float* const* images, size_t image_size, float cell_size,
                                    const svdi_options* opts, svdi_plan_info* infos);

/* Batched imaging of snapshots of mixed image sizes and cells (no counterpart in the reference;
 * the batching form): snapshot i is imaged into images[i], image_sizes[i]^2 floats, with cell
 * cell_sizes[i].  The snapshots are dealt to the concurrent lanes largest first.  wait != 0:
 * returns when every image is done (as svdi_image_snapshot_batch); wait == 0: returns once the
 * work is enqueued (as svdi_image_snapshot_batch_async). */
int svdi_image_snapshots(svdi_context* ctx, const svdi_snapshot* snapshots, size_t count,
                         float* const* images, const size_t* image_sizes, const float* cell_sizes,
                         const svdi_options* opts, svdi_plan_info* infos, int wait);

/* svdi_image_snapshots on plans earlier calls reported (no counterpart in the reference; the
 * batching form's graph): plans[i] is snapshot i's svdi_plan_info from a call on the same uvw,
 * frequencies, flags, image size and cell (as svdi_image_snapshot_planned).  Every snapshot's
 * work is enqueued on the concurrent lanes without any wait on the host; the lanes fork from the
 * context stream and the context stream joins them, so the call returns at once and the images
 * are complete when that stream reaches this point.  Nothing in the call allocates, waits or
 * records a memset when the lanes and their workspaces already hold the batch's shapes (an
 * earlier batched call of the same snapshots and sizes on the context), so the whole batch can be
 * captured into one graph on the context's stream (svdi_set_stream); under capture a context
 * without those lanes or shapes fails with SVDI_ERR_INVALID_ARGUMENT. */
int svdi_image_snapshots_planned(svdi_context* ctx, const svdi_snapshot* snapshots, size_t count,
                                 float* const* images, const size_t* image_sizes, const float* cell_sizes,
                                 const svdi_options* opts, const svdi_plan_info* plans);

/* Work shape of the MFMA window gridder (no counterpart in the reference, whose gridder has no
 * shape to tune; the batching form picks it per image size with svdi_autotune).  The gridder
 * contracts 16 x 16 windows of the grid; a window's records are contracted by one wavefront
 * when there are at most coop_min of them, else in pieces of at most coop_cap records each
 * taken by a whole workgroup.  Records are binned by row classes of 2^class_log2 rows.  Zero
 * fields select the library's defaults (coop_min 1024, coop_cap 24576, class_log2 2 when its
 * histogram fits in LDS, else 3).  The shape changes how the FP32 sums are ordered, not what
 * is summed: images agree to rounding. */
typedef struct svdi_tuning {
    uint32_t coop_min;          /* 0, or 32 .. 65536 */
    uint32_t coop_cap;          /* 0, or 4096 .. 16777216 */
    int32_t class_log2;         /* 0, 2 or 3 */
    float grid_ms;              /* svdi_autotune: the grid stage's time with this shape, ms */
    float default_ms;           /* svdi_autotune: the grid stage's time with the defaults, ms */
} svdi_tuning;

/* The shape the context's imaging calls use for grids of grid_size (= ceil(3 image_size / 2));
 * grid_size 0 sets the shape for every size without one of its own.  tuning NULL removes the
 * entry.  Batched calls pass the context's shapes to their lanes. */
int svdi_set_tuning(svdi_context* ctx, size_t grid_size, const svdi_tuning* tuning);
/* The shape in use for grid_size (all fields 0 when the defaults are). */
int svdi_get_tuning(svdi_context* ctx, size_t grid_size, svdi_tuning* tuning);

/* Autotuner (no counterpart in the reference): images `snapshot` at image_size / cell_size with
 * the MFMA gridder under each candidate shape (coop_min 256 .. 2048, coop_cap 12288 .. 49152, both
 * class heights), `repeats` times each (>= 1), times the grid stage with HIP events, and keeps the
 * fastest for this grid size on the context (svdi_set_tuning); *best receives it with its time
 * and the defaults' time.  K, the slabs and ES support stay the reference's: the search changes
 * only the work decomposition.  Allocates an image-sized scratch buffer for the trial images. */
int svdi_autotune(svdi_context* ctx, const svdi_snapshot* snapshot, size_t image_size, float cell_size,
                  const svdi_options* opts, int repeats, svdi_tuning* best);

/* Streaming request service (no counterpart in the reference, which images one snapshot per
 * process; the streaming form): snapshot requests submitted over time are admitted against a
 * device-memory budget, grouped by (image size, cell) into asynchronous batched calls and timed
 * from submission to image-ready, by a request loop on a thread of the library's own.
 *
 * A request costs its snapshot's device bytes (as the caller states them) plus its image; it is
 * admitted in submission order while the context's workspace plus the bytes already admitted
 * plus its cost fit budget_bytes.  Size the workspace before starting (batched calls of every
 * shape to be served, max_batch snapshots each), since it only grows: a request that cannot fit
 * beside the workspace is refused at submission, and one that stops fitting because the
 * workspace grew fails with SVDI_ERR_OUT_OF_MEMORY.  Admitted requests wait per key; while
 * fewer than max_inflight calls are in flight, the key of the oldest waiting request is imaged
 * with up to max_batch of its requests in one svdi_image_snapshots call on the context stream,
 * and an event behind it tells the loop when the images are ready.  The service uses the
 * context from its own thread: the caller must not image on the context meanwhile, and must
 * keep each request's snapshot and image allocated until its result is collected. */
typedef struct svdi_service svdi_service;
typedef struct svdi_service_config {
    size_t budget_bytes;        /* device bytes for the workspace and the admitted requests */
    uint32_t max_batch;         /* requests per batched call, >= 1 */
    uint32_t max_inflight;      /* batched calls in flight, >= 1 */
    svdi_options options;       /* imaging options of every call */
} svdi_service_config;
typedef struct svdi_service_result {
    uint64_t id;                /* the id given at submission */
    int32_t status;             /* svdi_status of the request's batched call */
    uint32_t batch;             /* requests in that call */
    uint64_t call;              /* that call's sequence number (0, 1, ...) */
    double queue_ms;            /* submission to launch */
    double latency_ms;          /* submission to the loop seeing the image ready */
} svdi_service_result;
typedef struct svdi_service_stats_info {
    size_t outstanding;         /* submitted, result not yet produced */
    size_t waiting;             /* submitted, not admitted */
    size_t admitted;            /* admitted, not launched */
    size_t inflight_calls;      /* batched calls in flight */
    size_t held_bytes;          /* bytes of admitted requests */
    size_t workspace_bytes;     /* the context's workspace as last seen */
} svdi_service_stats_info;

int svdi_service_create(svdi_context* ctx, const svdi_service_config* config, svdi_service** service);
/* Waits for every submitted request, then stops the loop and frees the service. */
int svdi_service_destroy(svdi_service* service);
/* image: device memory for image_size^2 floats, written by the call that images the request. */
int svdi_service_submit(svdi_service* service, uint64_t id, const svdi_snapshot* snapshot,
                        size_t snapshot_bytes, float* image, size_t image_size, float cell_size);
/* Up to `capacity` finished requests' results, oldest first; wait_ms > 0 waits up to that long
 * for one when none is ready and requests are outstanding. */
int svdi_service_collect(svdi_service* service, svdi_service_result* results, size_t capacity, size_t* count,
                         int wait_ms);
int svdi_service_stats(svdi_service* service, svdi_service_stats_info* info);

/* FIpipe (archive/FIPkernels.cu:679-911) on host buffers: Visreal/Visimag (S each), Bin
 * (S x 3), Vin (3 x 3), dirty_image (image_size^2, caller-allocated host memory). */
int svdi_fipipe_host(svdi_context* ctx, const float* vis_re, const float* vis_im,
                     const float* bin, const float* vin, float* dirty_image,
                     size_t num_baselines, size_t image_size, float cell_size,
                     const svdi_options* opts, svdi_plan_info* info);

/* The library's batched 2D inverse FFT on its own (the transform that replaces
 * cufftExecC2C(..., CUFFT_INVERSE), SVDIntegratedImager.cu:1658-1676,1713), for cross-checking
 * against another FFT: planes (num_planes x grid_size x grid_size complex64, device, row-major)
 * -> out (num_planes x rows x cols complex64, device) = rows [row_lo, row_lo + rows) and columns
 * [col_lo, col_lo + cols) of each unnormalised inverse transform, sum over (a, b) of
 * plane[a][b] e^{+2 pi i (a r + b c) / grid_size}.  The imager keeps the central image_size
 * rows and columns of an even grid, the whole plane of an odd one.  Returns when done. */
int svdi_fft2_inverse(svdi_context* ctx, const void* planes, size_t num_planes, size_t grid_size,
                      size_t row_lo, size_t rows, size_t col_lo, size_t cols, void* out);

/* One TIME snapshot read on the host from a MeasurementSet main table whose columns were
 * exported as NumPy .npy files (python-casacore's getcol arrays saved with numpy.save): the
 * integrated CLI's reader (read_measurement_set_snapshot, group_rows_by_time and
 * read_measurement_set_rows, SVDIntegratedImager.cu:180-378) with casacore's table layer
 * replaced by one file per column.  column_dir holds TIME.npy, UVW.npy, DATA.npy, FLAG.npy,
 * optionally WEIGHT_SPECTRUM.npy, SPECTRAL_WINDOW/CHAN_FREQ.npy and optionally
 * OBSERVATION/ARRAY_CENTER.npy.  Rows are grouped by exact TIME value in ascending order;
 * polarisations 0 and 3 are taken on the reference's row-axis layout rule.  `data` holds host
 * pointers owned by the library (weight0 / weight3 NULL without WEIGHT_SPECTRUM); free with
 * svdi_ms_free_snapshot.  Fails with SVDI_ERR_IO, svdi_last_error() giving the reference's
 * message where it has one ("Snapshot_Index is outside available TIME snapshots", "DATA has
 * fewer than 4 polarisations in ..."). */
typedef struct svdi_ms_snapshot {
    svdi_snapshot data;         /* host pointers */
    size_t index;               /* the snapshot's place in TIME order */
    size_t total_snapshots;     /* distinct TIME values of the table */
    double time;                /* its TIME value */
    void* internal;             /* the library's storage; do not touch */
} svdi_ms_snapshot;
int svdi_ms_read_snapshot(const char* column_dir, size_t snapshot_index, svdi_ms_snapshot* out);
void svdi_ms_free_snapshot(svdi_ms_snapshot* snapshot);

/* svdi_image_snapshot on a snapshot in host memory (svdi_snapshot with host pointers, e.g.
 * from svdi_ms_read_snapshot): the arrays are copied into the context's workspace and the image
 * (image_size^2 floats, caller-allocated host memory) is copied back.  This is
 * preprocess_measurement_set_gpu + FIpipeDevice of the integrated CLI
 * (SVDIntegratedImager.cu:635-853,1498-1761, main :1773-1853), integrated semantics by
 * default. */
int svdi_image_host_snapshot(svdi_context* ctx, const svdi_snapshot* host_snapshot, float* image,
                             size_t image_size, float cell_size, const svdi_options* opts,
                             svdi_plan_info* info);

/* FITS primary-HDU images, BITPIX -32 (write_fits_2d, SVDIntegratedImager.cu:153-176; the
 * archive's read_fits_image / write_fits_image, pipe_interface.cpp:12-137).  write replaces
 * an existing file.  read returns 1-D (naxes[1] = 1) or 2-D images as float; free with
 * svdi_free. */
int svdi_fits_write(const char* path, const float* data, long naxis1, long naxis2);
int svdi_fits_read(const char* path, float** data, long naxes[2]);
void svdi_free(void* p);

#ifdef __cplusplus
}  /* extern "C" */

/* archive/pipe_interface.cpp:8 — C++ linkage, mangled _Z6FIpipePfS_S_S_S_mmf.  Host buffers;
 * prints the archive's Chebyshev range check and "Time elapsed" like FIPkernels.cu:733-757,888;
 * returns 0 on success, nonzero on failure. */
int FIpipe(float* Visreal, float* Visimag, float* Bin, float* Vin, float* dirty_image,
           size_t num_baselines, size_t image_size, float cell_size);
#endif

#endif /* SVDIMAGER_H */
