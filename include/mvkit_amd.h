/* Extensions of the MI355X NCC build that have no counterpart in mvkit's C
 * ABI: a device-resident find for callers whose images already live in HBM
 * (benchmarks, GPU pipelines), a find that also returns every intermediate
 * stage (for stage-by-stage parity checks), and read-only model
 * introspection.  Results use the reference's internal pose convention
 * (template centroid, no HALCON half-pixel shift). */
#ifndef MVKIT_AMD_EXT_H
#define MVKIT_AMD_EXT_H

#include "mv_ncc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Coarse correlation forms (same values as NCCHIP_FORM_*).  AUTO: split-BF16
 * on byte images, FP32 on uint2 images (what mv_find_ncc_model runs).  A
 * split-BF16 or FP8-screened search whose top level or coarse bounding box
 * does not fit the matrix-core kernel runs the FP32 form instead
 * (MvAmdFindStats.coarse_form reports the form that ran). */
#define MVAMD_FORM_AUTO (-1)
#define MVAMD_FORM_FP32 0
#define MVAMD_FORM_BF16X3 1
#define MVAMD_FORM_FP64 2
#define MVAMD_FORM_FP8_SCREEN 3

typedef struct MvAmdFindOptions {
  double angle_start, angle_extent, min_score, max_overlap;
  int64_t num_matches;
  int32_t subpixel;
  int32_t highest_level, lowest_level;  /* 1-based; highest 0 = all */
  int32_t form;                          /* MVAMD_FORM_* */
  double screen;                         /* FP8 screen threshold (form FP8_SCREEN; -inf: 0.55) */
  int32_t profile;                       /* bit s: time stage s (MVAMD_STAGE_*) with HIP events;
                                            MVAMD_PROFILE_ALL times every stage */
  int32_t vector_tracking;               /* 1: track every level with the FP64 vector kernel
                                            (0: matrix cores where they apply; same results) */
  int32_t pixel_type;                    /* MVAMD_PIXEL_*: pixel type of the device images
                                            (pitches in bytes) */
  int64_t max_candidates;                /* coarse candidate budget: the reference's internal
                                            FindOptions::max_candidates (model.hpp:21), which its
                                            C ABI leaves at 10 000; 0 selects 10 000 */
  int32_t no_graph;                      /* 0: the device part of a find whose model, options, frame
                                            geometry and workspace repeat is replayed as a captured
                                            HIP graph; 1: always launched kernel by kernel */
  int32_t coarse_tile;                   /* tile shape of the split-BF16 coarse kernel:
                                            MVAMD_COARSE_TILE_AUTO (0) the model's autotuned shape,
                                            MVAMD_COARSE_TILE_DEFAULT (-1), or MVAMD_COARSE_TILE(...) */
} MvAmdFindOptions;

/* Tile shapes of the split-BF16 coarse kernel.  A wave computes `rows`
 * output rows (4 or 8) x 32 anchors against 32 rotations; a block has
 * col_waves x row_waves waves (each 1, 2 or 4, product <= 4) and runs `tiles`
 * angle tiles of 32 rotations (1..8) over one image tile staged in LDS.
 * MVAMD_COARSE_TILE_BOUND adds the bound group: 0 bounds each window's score
 * by the support common to every rotation of the model, 8, 16 or 32 by the
 * support common to its group of that many consecutive rotations (a tighter
 * bound computed per group, so fewer windows need exact moments; not with
 * MVAMD_FORM_FP8_SCREEN).
 * Every shape gives bit-identical planes.  With MVAMD_COARSE_TILE_AUTO the
 * first find of a kind of search (level, pixel type, coarse angles, template
 * extent, screen, frame count and size class) times the candidate shapes on
 * its own frames and keeps the fastest in the model for later finds
 * (mvamd_get_ncc_model_tuning / mvamd_set_ncc_model_tuning read and restore
 * those choices). */
#define MVAMD_COARSE_TILE_AUTO 0
#define MVAMD_COARSE_TILE_DEFAULT (-1)
#define MVAMD_COARSE_TILE(rows, col_waves, row_waves, tiles) \
  ((rows) | ((col_waves) << 4) | ((row_waves) << 8) | ((tiles) << 12))
#define MVAMD_COARSE_TILE_BOUND(rows, col_waves, row_waves, tiles, bound) \
  (MVAMD_COARSE_TILE(rows, col_waves, row_waves, tiles) |                  \
   (((bound) == 8 ? 1 : (bound) == 16 ? 2 : (bound) == 32 ? 3 : 0) << 16))

/* Pixel types of device-resident search images ("byte" and "uint2" in the
 * reference's image objects, c_api/ncc_matching_api.cpp:167-168). */
#define MVAMD_PIXEL_BYTE 0
#define MVAMD_PIXEL_UINT2 2

/* Kernel stages reported in MvAmdFindStats.kernel_us. */
#define MVAMD_STAGE_PYRAMID 0
#define MVAMD_STAGE_PREFIX 1
#define MVAMD_STAGE_COARSE 2
#define MVAMD_STAGE_PEAKS 3
#define MVAMD_STAGE_TRACK 4
#define MVAMD_STAGE_FULL 5
#define MVAMD_STAGE_REFINE 6
#define MVAMD_STAGE_TRACK_MFMA 7   /* tracking levels on matrix cores (MVAMD_STAGE_TRACK: FP64 vector kernel) */
#define MVAMD_STAGE_COUNT 8
#define MVAMD_PROFILE_ALL 0xff

typedef struct MvAmdFindStats {
  int32_t n_coarse, n_prepared, top, finest, angle_stride;
  int32_t storage[4];
  double t_coarse_us, t_track_us, t_refine_us, t_total_us;
  double kernel_us[MVAMD_STAGE_COUNT];   /* device time of each profiled stage */
  double coarse_flops;                   /* algorithmic FLOPs of the coarse correlation */
  double track_flops;                    /* algorithmic FLOPs of tracking (when its stage is profiled) */
  double track_mfma_flops;               /* the part of track_flops tracked on matrix cores */
  int32_t coarse_form;                   /* MVAMD_FORM_* the coarse search ran (-1: none ran) */
  int32_t coarse_tile;                   /* MVAMD_COARSE_TILE(...) shape of the split-BF16 coarse
                                            kernel that ran (0: it did not run) */
} MvAmdFindStats;

/* Find on a u8 (or, with pixel_type MVAMD_PIXEL_UINT2, u16) image given as
 * a HIP device pointer on the calling thread's current device.  results: cap x 4 doubles (row, col, angle, score).
 * Replaces, for device-resident inputs, mv_find_ncc_model
 * (mvkit include/mvkit/c/mv_ncc_matching.h:71-74) minus the output shift. */
MV_API MvStatus mvamd_find_ncc_model_device(MvTuple ModelID, const uint8_t* d_image, int64_t width, int64_t height,
                                            int64_t pitch, const MvAmdFindOptions* options, double* results,
                                            size_t cap, size_t* count, MvAmdFindStats* stats);

/* Find on an image object (host pixels, u8 or u16, with its reduced domain
 * if it has one) with the extension options (correlation form, profiling;
 * pixel_type is taken from the image).  results: cap x 4 doubles (row, col,
 * angle, score) in the internal (centroid) convention.  Replaces
 * mv_find_ncc_model (mvkit include/mvkit/c/mv_ncc_matching.h:71-74) minus
 * the output shift, for callers that choose the correlation form. */
MV_API MvStatus mvamd_find_ncc_model_image(MvObject Image, MvTuple ModelID, const MvAmdFindOptions* options,
                                           double* results, size_t cap, size_t* count, MvAmdFindStats* stats);

/* find_ncc_models on a u8 or u16 image given as a HIP device pointer.  options holds
 * one entry or one per model (n_options); the selection is what
 * mv_find_ncc_models derives from one-value controls: selection_matches
 * (-1: per model) and selection_overlap (-1: per model, same-model pairs
 * only).  Matches go to results as cap x 5 doubles (row, col, angle, score,
 * model index) in the internal (centroid) convention; their number to count.
 * Replaces, for device-resident inputs, mv_find_ncc_models
 * (mvkit include/mvkit/c/mv_ncc_matching.h:119-123) minus the output shift. */
MV_API MvStatus mvamd_find_ncc_models_device(MvTuple ModelIDs, const uint8_t* d_image, int64_t width, int64_t height,
                                             int64_t pitch, const MvAmdFindOptions* options, size_t n_options,
                                             int64_t selection_matches, double selection_overlap, double* results,
                                             size_t cap, size_t* count);

/* find_ncc_models over n_images u8 (or u16) images given as HIP device
 * pointers, with the models, options and selection of
 * mvamd_find_ncc_models_device; image k's matches go to results + 5*cap*k
 * (row, col, angle, score, model index) and their number to counts[k], each
 * image's result the one mvamd_find_ncc_models_device returns for it.  One
 * query per distinct (model, options) covers all images (one launch per
 * stage); coarse_queue != 0 runs the split-BF16 coarse searches of every
 * query in one persistent launch per kernel instantiation fed by an (image,
 * angle tiles, block) work queue, the items of the queries that cost most
 * first (0: one launch per query).  At most 256 images.  stats
 * (optional) receives the call's totals.
 * Replaces, for many device-resident images, mv_find_ncc_models
 * (mvkit include/mvkit/c/mv_ncc_matching.h:119-123) per image. */
typedef struct MvAmdBatchStats {
  double coarse_queue_us;   /* device time of the persistent coarse launches (HIP events) */
  double coarse_flops;      /* algorithmic FLOPs of every query's coarse search */
  double total_us;          /* host time of the call */
  double tune_us;           /* host time of the call spent tuning coarse tile shapes */
  int32_t queries;          /* distinct (model, options) queries */
  int32_t queue_queries;    /* those searched in the persistent launch */
  int32_t candidates;       /* full-scored candidates over all queries and images */
  int32_t tuned;            /* queries whose tile shape this call tuned */
  int32_t refined;          /* candidates refined (the walks' speculative rounds included) */
  int32_t refine_rounds;    /* refinement rounds (one launch and one synchronisation each) */
} MvAmdBatchStats;
MV_API MvStatus mvamd_find_ncc_models_batch(MvTuple ModelIDs, const uint8_t* const* d_images, const int64_t* widths,
                                            const int64_t* heights, const int64_t* pitches, size_t n_images,
                                            const MvAmdFindOptions* options, size_t n_options,
                                            int64_t selection_matches, double selection_overlap, int32_t coarse_queue,
                                            double* results, size_t cap, size_t* counts, MvAmdBatchStats* stats);

/* A request loop on the calling thread's current device: requests (one host
 * image, 1..n models with find_ncc_models' options and selection) queue up,
 * are admitted while the estimated device working sets of the requests in
 * flight stay within hbm_budget bytes (the oldest request first; always when
 * none is in flight), and run in micro-batches of up to max_batch requests
 * by `workers` threads whose device work overlaps on the GPU; models stay
 * resident on the device.  Each request's result (row, col, angle, score,
 * model index; internal convention) is the one mvamd_find_ncc_models_device
 * returns for it alone.  Zero fields select the defaults (2 workers, 8
 * requests per batch, 16 GiB, 4096 waiting requests).  Submit and wait may
 * be called from any threads; destroy waits for the requests in flight and
 * must not race with submit or wait.  Replaces, for a serving loop,
 * concurrent mv_find_ncc_models calls on shared models
 * (docs/runtime.md:31). */
typedef struct MvAmdServerConfig {
  int32_t workers, max_batch;
  int64_t hbm_budget;
  int64_t queue_limit;
  int32_t no_coarse_queue;  /* 1: one coarse launch per query instead of a persistent launch per batch */
  int32_t reserved;
} MvAmdServerConfig;
typedef struct MvAmdRequestTimes {
  double admitted_us, started_us, finished_us;  /* after submission */
  double bytes;                                 /* estimated working set it was admitted with */
  int32_t batch_size, worker;
} MvAmdRequestTimes;
MV_API MvStatus mvamd_server_create(const MvAmdServerConfig* config, void** server);
MV_API MvStatus mvamd_server_destroy(void* server);
MV_API MvStatus mvamd_server_submit(void* server, const uint8_t* image, int64_t width, int64_t height, int64_t pitch,
                                    MvTuple ModelIDs, const MvAmdFindOptions* options, size_t n_options,
                                    int64_t selection_matches, double selection_overlap, uint64_t* request);
MV_API MvStatus mvamd_server_wait(void* server, uint64_t request, double* results, size_t cap, size_t* count,
                                  MvAmdRequestTimes* times);

/* Finds the model in n u8 (or u16) images given as HIP device pointers, all
 * with the same options, stage by stage over the whole batch (one launch per stage for
 * the candidates of every frame).  Frame k's matches go to
 * results + 4*cap*k (row, col, angle, score) and their number to counts[k];
 * each frame's result is the one mvamd_find_ncc_model_device returns for it.
 * stats (optional) receives the batch totals of frame 0's trace. */
MV_API MvStatus mvamd_find_ncc_model_batch(MvTuple ModelID, const uint8_t* const* d_images, const int64_t* widths,
                                           const int64_t* heights, const int64_t* pitches, size_t n,
                                           const MvAmdFindOptions* options, double* results, size_t cap,
                                           size_t* counts, MvAmdFindStats* stats);

/* The two halves of mvamd_find_ncc_model_batch, for overlapping consecutive
 * batches on the calling thread's current device: submit enqueues the
 * device-resident part of the find (at most 256 frames) and returns at once
 * with a ticket; wait completes it (refinement rounds, results as in
 * mvamd_find_ncc_model_batch).  At most two batches may be pending per
 * thread and device; a batch submitted after another overlaps the other's
 * refinement and host work.  The images must stay valid until the wait
 * returns; the model stays alive as long as the batch is pending.
 * with_stats != 0 fills stats at the wait. */
MV_API MvStatus mvamd_find_ncc_model_batch_submit(MvTuple ModelID, const uint8_t* const* d_images,
                                                  const int64_t* widths, const int64_t* heights, const int64_t* pitches,
                                                  size_t n, const MvAmdFindOptions* options, int32_t with_stats,
                                                  uint64_t* ticket);
MV_API MvStatus mvamd_find_ncc_model_batch_wait(uint64_t ticket, double* results, size_t cap, size_t* counts,
                                                MvAmdFindStats* stats);

/* Find on a host image (u8, or u16 with options->pixel_type MVAMD_PIXEL_UINT2;
 * pitch in bytes) returning the stages: coarse peaks (cap_coarse x 4),
 * full-scored candidates in score order (cap_prepared x 4) and, per such
 * candidate, the refined pose and verdict (cap_prepared x 5). */
MV_API MvStatus mvamd_find_ncc_model_trace(MvTuple ModelID, const uint8_t* image, int64_t width, int64_t height,
                                           int64_t pitch, const MvAmdFindOptions* options, double* results,
                                           size_t cap, size_t* count, double* coarse, size_t cap_coarse,
                                           size_t* n_coarse, double* prepared, double* refined, size_t cap_prepared,
                                           size_t* n_prepared, MvAmdFindStats* stats);

/* mv_create_ncc_model (mvkit include/mvkit/c/mv_ncc_matching.h:30-33) with a
 * choice of where the rotated rasters are built: MVAMD_CREATE_HOST (host
 * FP64), MVAMD_CREATE_DEVICE (the calling thread's current HIP device; error
 * without one) or MVAMD_CREATE_AUTO (the device when one is present, what
 * mv_create_ncc_model does).  Both builders give bit-identical models. */
#define MVAMD_CREATE_AUTO 0
#define MVAMD_CREATE_HOST 1
#define MVAMD_CREATE_DEVICE 2
MV_API MvStatus mvamd_create_ncc_model(MvObject Template, MvTuple NumLevels, MvTuple AngleStart, MvTuple AngleExtent,
                                       MvTuple AngleStep, MvTuple Metric, int32_t builder, MvTuple* ModelID);

/* The model's autotuned coarse tile shapes as text, one line per kind of
 * search: "<key> : rows=R col_waves=C row_waves=W tiles=T us=U default_us=D"
 * (key: level, pixel kind, coarse angles, union box rows and 16-column
 * chunks, screen, then the frame count and top-level size classes of a
 * search launched on its own, or "queue" for one run in the shared
 * persistent launch of mvamd_find_ncc_models_batch and the server; us and
 * default_us the sample launch's time with the chosen and the default
 * shape).  text
 * receives at most cap bytes including the terminating zero; *length the
 * full length without it (call with cap 0 to size the buffer).  No
 * counterpart in the reference (its CPU kernels have no tile shapes). */
MV_API MvStatus mvamd_get_ncc_model_tuning(MvTuple ModelID, char* text, size_t cap, size_t* length);
/* Replaces the model's tuned shapes with the lines of `text` (as
 * mvamd_get_ncc_model_tuning writes them; empty text clears them, so the next
 * find of each kind tunes again).  Lines whose shape is invalid are rejected
 * (INVALID_ARGUMENT) and nothing is changed. */
MV_API MvStatus mvamd_set_ncc_model_tuning(MvTuple ModelID, const char* text);

/* Model introspection.  info: levels, angles per level, angle step,
 * origin row, origin column, sample count. */
MV_API MvStatus mvamd_model_info(MvTuple ModelID, double* info);
/* Raster of (level, angle index), level -1 = the unrotated original patch.
 * header: n, angle, orow, ocol, rmin, rmax, cmin, cmax, weight sum, #runs,
 * #weight runs.  rows/cols/weights receive up to cap samples. */
MV_API MvStatus mvamd_model_raster(MvTuple ModelID, int32_t level, int32_t angle, double* header, int32_t* rows,
                                   int32_t* cols, double* weights, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
