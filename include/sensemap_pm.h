/* sensemap_pm.h -- C-ABI of the MI355X (gfx950) PatchMatch multi-view-stereo
 * engine.  This is the drop-in boundary for SenseMap's PatchMatch stage:
 *
 *   reference (src/mvs/patch_match_cuda.h:154-203)        this ABI
 *   ------------------------------------------------------------------------
 *   PatchMatchACMMCuda(options, problem)   (ctor)      pm_problem_create
 *   bool PatchMatchACMMCuda::Run()                     pm_problem_run
 *   DepthMap GetDepthMap() const                       pm_problem_get_depth_map
 *   NormalMap GetNormalMap() const                     pm_problem_get_normal_map
 *   Mat<float> GetConfMap() const                      pm_problem_get_conf_map
 *   cost_map_ (no getter in the reference)             pm_problem_get_cost_map
 *   view_sel_map_ (no getter in the reference)         pm_problem_get_view_sel
 *   ~PatchMatchACMMCuda()                              pm_problem_destroy
 *   PatchMatchOptions defaults / Check()               pm_options_init / pm_options_check
 *     (src/controllers/patch_match_options.h:42-304, patch_match_options.cc:75-101)
 *   SetBestCudaDevice(gpu_index) (src/util/cuda.cc:36-61)  pm_engine_create(device)
 *
 * Plain C: fixed-width scalars and pointers only.  Every call returns a
 * pm_status and never exits or aborts; pm_last_error() returns the message of
 * the calling thread's last failure.  Inputs are borrowed for the duration of
 * the call that receives them; outputs go to caller-allocated buffers.
 */
#ifndef SENSEMAP_PM_H_
#define SENSEMAP_PM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_ABI_VERSION 6
#define PM_MAX_NUM_SRC 32      /* the reference caps at MAX_NUM_SRC_IMAGE = 8 */
#define PM_MAX_WINDOW_RADIUS 32
#define PM_POSE_STRIDE 44      /* floats of packed pose per source view */

typedef enum pm_status {
  PM_OK = 0,
  PM_ERR_INVALID_ARG = 1,
  PM_ERR_OUT_OF_MEMORY = 2,
  PM_ERR_DEVICE = 3,
  PM_ERR_TOO_MANY_SOURCES = 4,
  PM_ERR_NOT_READY = 5,
  PM_ERR_OVER_BUDGET = 6,
  PM_ERR_INTERNAL = 7
} pm_status;

enum { PM_STORAGE_U8 = 0, PM_STORAGE_F16 = 1 };
enum { PM_ACCUM_PLAIN = 0, PM_ACCUM_COMPENSATED = 1 };
enum { PM_RCP_EXACT = 0, PM_RCP_FAST = 1 };

/* Engine-relevant fields of PatchMatchOptions, same names and defaults
 * (src/controllers/patch_match_options.h:42-304).  Fields the engine does not
 * read (fusion, meshing, lidar, cache sizes, ...) are not mirrored. */
typedef struct pm_options {
  int32_t window_radius;              /* 5 */
  int32_t window_step;                /* 2 */
  float sigma_spatial;                /* = window_radius (5) */
  float sigma_color;                  /* 0.2 */
  int32_t num_samples;                /* 15 */
  int32_t num_iterations;             /* 5 */
  int32_t geom_consistency;           /* 1 */
  float geom_consistency_regularizer; /* 0.3 */
  float geom_consistency_max_cost;    /* 3.0 */
  int32_t est_curvature;              /* 0: ComputeCurvatureMap runs when geom_consistency && est_curvature (:71) */
  int32_t random_optimization;        /* 1 */
  int32_t local_optimization;         /* 1 */
  float random_depth_ratio;           /* 0.004 */
  float random_angle1_range;          /* 20 deg, in radians */
  float random_angle2_range;          /* 12 deg, in radians */
  float init_ncc_matching_cost;       /* 0.8 */
  float max_ncc_matching_cost;        /* 1.2 */
  float alpha;                        /* 90 */
  float beta;                         /* 0.3 */
  int32_t thk;                        /* 4 (1..4) */
  int32_t num_good_hypothesis;        /* 2 */
  int32_t num_bad_hypothesis;         /* 3 */
  int32_t init_depth_random;          /* = !geom_consistency */
  int32_t filter;                     /* 1 */
  int32_t conf_filter;                /* 0: FilterDepthMap runs when filter && conf_filter */
  int32_t median_filter;              /* 0: ComputeGeomGradientMap + Black/RedPixelFilter after the iterations (:248) */
  float conf_threshold;               /* 0.5 */
  float depth_diff_threshold;         /* 0.02 */
  float filter_geom_consistency_max_cost;  /* 1.0: RunFilter's per-source threshold (patch_match_options.h:139) */
  int32_t filter_min_num_consistent;  /* 1: RunFilter keeps a depth with this many consistent sources (:135) */
  /* Priors (patch_match_options.h; read at patch_match_cuda.cu:377, 603, 943-947): 0 only.  has_prior_depth
   * belongs to the RGB-D path; propagate_depth makes the reference's sweep read prior normal and weight maps
   * it allocates as 0x0 (patch_match_cuda.cu:589-591, kernel_functions.cu:2191-2199), so it defines no result.
   * pm_options_check rejects either with PM_ERR_INVALID_ARG rather than ignore it. */
  int32_t propagate_depth;            /* 0 */
  int32_t has_prior_depth;            /* 0 */
  /* ---- extensions (no counterpart in the reference) ---- */
  int32_t init_samples;               /* 1 = reference semantics; K > 1 keeps the best of K random inits */
  int32_t compat_lo_geo_bug;          /* 1 = reproduce kernel_functions.cu:2924 (ref_invK passed as K) */
  int32_t image_storage;              /* source images in HBM: PM_STORAGE_F16 (default; the 8-bit values exactly,
                                         fewer instructions per sample) or PM_STORAGE_U8 (half the bytes); same maps */
  int32_t accum_mode;                 /* PM_ACCUM_PLAIN or PM_ACCUM_COMPENSATED (NCC moments) */
  int32_t rcp_mode;                   /* PM_RCP_EXACT (bitwise with the CPU transcription) or PM_RCP_FAST */
  int32_t reserved0;
  int32_t reserved1;
  uint64_t seed;                      /* Philox key; streams are keyed by (seed, view_id, pixel, pass) */
} pm_options;

/* Pinhole camera: K = {fx, cx, fy, cy}; world->camera rotation R (row-major)
 * and translation T, as mvs::Image::GetK/GetR/GetT (src/mvs/image.h). */
typedef struct pm_camera {
  float K[4];
  float R[9];
  float T[3];
} pm_camera;

/* 8-bit greyscale image, row 0 at the top (src/util/bitmap.cc:201-205). */
typedef struct pm_image {
  const uint8_t* data;
  int32_t width;
  int32_t height;
  int32_t stride;  /* bytes between rows, >= width */
  /* > 0: a serving loop may keep this image on the device under this key
   * (images given the same key must be identical; a key the server holds
   * for an image of another size fails the job with PM_ERR_INVALID_ARG);
   * <= 0: never cached.  Ignored outside pm_server_*. */
  int32_t cache_id;
} pm_image;

/* One reference view with its source views: the fields of Problem
 * (src/controllers/patch_match_options.h:306-353) the engine reads, with the
 * images, cameras and maps given directly instead of by index into
 * controller-owned vectors. */
typedef struct pm_view_desc {
  uint32_t view_id;                 /* Problem::ref_image_idx; keys the random streams */
  int32_t num_src;                  /* Problem::src_image_idxs.size(), 1..PM_MAX_NUM_SRC */
  pm_image ref;
  pm_camera ref_cam;
  const pm_image* src;              /* num_src entries */
  const pm_camera* src_cams;        /* num_src entries */
  float depth_min;                  /* PatchMatchOptions::depth_min as set per problem */
  float depth_max;
  /* Geometric pass (geom_consistency): per-source depth maps, row-major,
   * src[i].width x src[i].height floats each (Problem::depth_maps). */
  const float* const* src_depth_maps;
  /* Initial state when !init_depth_random: depth W x H and normal 3 x W x H
   * planar (Problem::depth_maps / normal_maps at ref_image_idx). */
  const float* init_depth;
  const float* init_normal;
  /* Problem::mask_maps at ref_image_idx (patch_match_cuda.cu:113,173-196):
   * an 8-bit mask of the reference image's size; data NULL = no mask.  When
   * given, the run ends with MaskDepthMapKernel (:1040-1052). */
  pm_image ref_mask;
} pm_view_desc;

/* Caller-allocated outputs; a NULL pointer skips that map. */
typedef struct pm_outputs {
  float* depth;        /* W x H */
  float* normal;       /* 3 x W x H planar (x, y, z), the NormalMap layout */
  float* conf;         /* W x H */
  float* cost;         /* num_src x W x H planar: per-source cost of the winning hypothesis */
  uint32_t* view_sel;  /* W x H bitmask of selected source views */
  uint16_t* curvature; /* 2 x W x H planar {gradient, curvature}; written when geom_consistency && est_curvature */
} pm_outputs;

/* Timings of the last pm_problem_run, measured with hipEvents on the
 * engine's stream. */
typedef struct pm_run_stats {
  double total_ms;          /* whole Run() on the device */
  double sweep_ms;          /* sum over all checkerboard sweep launches */
  double local_opt_ms;      /* sum over local-optimisation launches */
  double init_ms;           /* init + init-conf launches */
  int32_t sweep_launches;
  int32_t local_opt_launches;
  /* When counting is enabled (pm_engine_set_counting): (hypothesis, source)
   * NCC evaluations, their window samples and geometric-consistency terms,
   * in total and for the sweep kernel alone. */
  uint64_t ncc_evaluations;
  uint64_t ncc_samples;
  uint64_t geo_evaluations;
  uint64_t sweep_ncc_evaluations;
  uint64_t sweep_ncc_samples;
  uint64_t sweep_geo_evaluations;
  /* the steps after the iterations: curvature map, median filter,
   * FilterDepthMap, mask */
  double post_ms;
} pm_run_stats;

typedef struct pm_engine pm_engine;
typedef struct pm_problem pm_problem;

/* Fill the PatchMatchOptions defaults (patch_match_options.h:42-304). */
pm_status pm_options_init(pm_options* options);
/* PatchMatchOptions::Check() (patch_match_options.cc:75-101) plus the limits
 * of this engine (thk 1..4, sources, radius).  PM_OK when valid. */
pm_status pm_options_check(const pm_options* options);

const char* pm_last_error(void);
int32_t pm_abi_version(void);

/* Select device `device` and create a HIP stream owned by the engine
 * (SetBestCudaDevice, src/util/cuda.cc:36-61). */
pm_status pm_engine_create(int32_t device, pm_engine** out);
/* Drop the caller's handle.  Problems, batches and servers created on the
 * engine each keep it alive, so they may be destroyed after it (in any
 * order, as a garbage collector finalises); the stream and the cached device
 * blocks are released with the last of them. */
void pm_engine_destroy(pm_engine* engine);
/* Turn on per-run counting of NCC evaluations and window samples
 * (calibration runs only; off by default). */
pm_status pm_engine_set_counting(pm_engine* engine, int32_t enable);

/* Upload the view, pack poses and allocate the workspace: the
 * PatchMatchACMMCuda constructor (patch_match_cuda.cu:422-432). */
pm_status pm_problem_create(pm_engine* engine, const pm_options* options,
                            const pm_view_desc* view, pm_problem** out);
/* Run the iteration schedule (patch_match_cuda.cu:748-1057); blocks until done. */
pm_status pm_problem_run(pm_problem* problem);
pm_status pm_problem_get_depth_map(const pm_problem* problem, float* out);
pm_status pm_problem_get_normal_map(const pm_problem* problem, float* out);
pm_status pm_problem_get_conf_map(const pm_problem* problem, float* out);
pm_status pm_problem_get_cost_map(const pm_problem* problem, float* out);
pm_status pm_problem_get_view_sel(const pm_problem* problem, uint32_t* out);
/* GetCurvatureMap (patch_match_cuda.h:162, patch_match_cuda.cu:444-446):
 * the 2 x W x H planar unsigned-short map of ComputeCurvatureMap
 * (kernel_functions.cu:1603-1651), slice 0 = min(10000 |Scharr gradient| of
 * the reference image in [0,1] units, 32767), slice 1 = 10000 x the median
 * angle (radians) between the pixel's normal and its 5x5 neighbours'.
 * PM_ERR_INVALID_ARG unless the problem was run with geom_consistency and
 * est_curvature (the reference allocates the map only then, :682-684). */
pm_status pm_problem_get_curvature_map(const pm_problem* problem, uint16_t* out);
pm_status pm_problem_get_stats(const pm_problem* problem, pm_run_stats* out);
/* Packed per-source pose block (PM_POSE_STRIDE floats per source) and the
 * reference {fx,cx,fy,cy, 1/fx,-cx/fx,1/fy,-cy/fy} (InitTransforms,
 * patch_match_cuda.cu:493-580). */
pm_status pm_problem_get_poses(const pm_problem* problem, float* poses, float ref_k8[8]);
void pm_problem_destroy(pm_problem* problem);

/* RunFilter (patch_match_cuda.cu:1074-1108): the view's given depth map
 * (init_depth, loaded as InitWorkspaceFilter does, :701-712) is checked
 * against the source depth maps with the geometric-consistency cost.
 * FilterDepthMapsKernel (kernel_functions.cu:3356-3450) zeroes a depth with
 * fewer than min(filter_min_num_consistent, num_src) sources within
 * filter_geom_consistency_max_cost; with deduplication,
 * DeduplicDepthMapsKernel (:3452-3490) zeroes a depth that any source
 * explains.  Needs a problem created with geom_consistency = 1 and
 * init_depth_random = 0; the result is its depth map. */
pm_status pm_problem_run_filter(pm_problem* problem, int32_t deduplication);

/* Device-to-device hand-off between passes (the reference goes through
 * host DepthMap/NormalMap caches, patch_match_controller.cc:4571-4595,
 * 4679-4680).  Both problems must live on the same engine.
 * pm_problem_init_from: use `from`'s current depth and normal maps as the
 *   initial state of `dst` (dst created with init_depth_random = 0, the
 *   geometric pass; InitWorkspace, patch_match_cuda.cu:648-665).
 * pm_problem_set_src_depth: replace the depth map of source `src_index` of
 *   `dst` (geometric pass input, patch_match_cuda.cu:318-368) with `from`'s
 *   current depth map; sizes must match. */
pm_status pm_problem_init_from(pm_problem* dst, const pm_problem* from);
pm_status pm_problem_set_src_depth(pm_problem* dst, int32_t src_index, const pm_problem* from);

/* ---- batched form (no counterpart in the reference) ----
 * Many reference views, of uneven size, source count and window shape, run
 * together: at every step each view advances by one phase of its Run()
 * schedule, and views whose phase uses the same kernel share one launch.
 * The reference runs one view per engine and serialises launches per host
 * thread (patch_match_controller.cc:3620-3632, patch_match_cuda.cu:841-880).
 * options and views hold num_views entries each; every view follows its own
 * options (iterations, window, geometric pass, ...).  Results are read per
 * view with pm_problem_get_* on the handle from pm_batch_get_view. */
typedef struct pm_batch pm_batch;
pm_status pm_batch_create(pm_engine* engine, int32_t num_views, const pm_options* options,
                          const pm_view_desc* views, pm_batch** out);
/* Capture the launch sequence into a hipGraph on the next run and replay it
 * afterwards (counting runs stay eager). */
pm_status pm_batch_set_graph(pm_batch* batch, int32_t enable);
pm_status pm_batch_run(pm_batch* batch);
pm_status pm_batch_get_view(const pm_batch* batch, int32_t index, pm_problem** view);
pm_status pm_batch_get_stats(const pm_batch* batch, pm_run_stats* out);
void pm_batch_destroy(pm_batch* batch);

/* ---- launch shapes and the autotuner (no counterpart in the reference) ----
 * Every colour-pass kernel runs persistent workgroups that claim tiles of
 * 2^tile_lw columns x 512 / 2^tile_lw rows (tile_lw 4, 5, 6: 8, 16 or 32
 * threads per tile row), with at most blocks_per_cu resident workgroups per
 * CU (0 = as many as fit).  The shape changes only the speed: every shape
 * gives the same maps, bit for bit.  An engine keeps one shape per class:
 * phase kernel, window (radius, step), image storage, reciprocal mode and
 * source-count bucket (N <= 8, N <= 16, N > 16); classes without one use
 * tile_lw 5 and blocks_per_cu 0. */
typedef enum pm_phase {
  PM_PHASE_INIT = 0,
  PM_PHASE_SWEEP = 1,
  PM_PHASE_LOCAL = 2,
  PM_PHASE_FILTER = 3,
  PM_PHASE_CURVATURE = 4,
  PM_PHASE_MASK = 5,
  PM_PHASE_GEOM_GRADIENT = 6,
  PM_PHASE_MEDIAN = 7
} pm_phase;

typedef struct pm_tuning {
  int32_t tile_lw;        /* 4..6 */
  int32_t blocks_per_cu;  /* >= 0 */
} pm_tuning;

typedef struct pm_tune_report {
  pm_tuning sweep;           /* chosen shapes */
  pm_tuning local_opt;
  double sweep_ms_default;   /* one sweep launch of the view, default shape */
  double sweep_ms_best;      /* ... chosen shape */
  double local_opt_ms_default;
  double local_opt_ms_best;
  int32_t candidates;        /* shapes timed per kernel */
  int32_t reserved0;
} pm_tune_report;

pm_status pm_engine_set_tuning(pm_engine* engine, int32_t phase, const pm_options* options, int32_t num_src,
                               const pm_tuning* tuning);
pm_status pm_engine_get_tuning(pm_engine* engine, int32_t phase, const pm_options* options, int32_t num_src,
                               pm_tuning* out);
/* Time every shape on this view's sweep and local-optimisation kernels and
 * keep the fastest in the engine for the view's classes.  Runs the view's
 * init and a few sweeps, so its maps are undefined until its next run. */
pm_status pm_problem_autotune(pm_problem* problem, pm_tune_report* report);
/* Keep the engine's shapes in a text file, one line per class, keyed by the
 * device architecture name; loading skips lines of other architectures. */
pm_status pm_engine_save_tuning(pm_engine* engine, const char* path);
pm_status pm_engine_load_tuning(pm_engine* engine, const char* path);
/* Before its first run, autotune every class of the batch that the engine
 * holds no shape for, on the class's largest view. */
pm_status pm_batch_set_autotune(pm_batch* batch, int32_t enable);

/* ---- serving loop (no counterpart in the reference) ----
 * A host thread owns the device side of a stream of depth-map jobs.  Jobs
 * are admitted first-come first-served while the device memory they need
 * fits in the HBM budget and fewer than max_active_views jobs are on the
 * device; source images (pm_image.cache_id > 0) stay on the device in an
 * LRU cache shared by all jobs and are evicted only when no job uses them.
 * Every step advances each admitted job by one phase of its Run() schedule,
 * and the phases of all jobs that share a kernel run in one launch
 * (continuous batching: jobs join and leave between steps).  A job's maps
 * do not depend on the other jobs: they equal, bit for bit, the maps of the
 * same job run alone with pm_problem_run on an engine with the same launch
 * shapes.  The server uses the engine's launch shapes (pm_engine_set_tuning,
 * pm_problem_autotune, pm_engine_load_tuning) and never tunes by itself.
 * The reference instead runs up to 3 engines per GPU from host threads, one
 * view each (patch_match_controller.cc:964-995, mvs/utils.h:33). */
typedef struct pm_server pm_server;

typedef struct pm_server_config {
  uint64_t hbm_budget_bytes;  /* device memory of the server: jobs, cache, workspace */
  int32_t max_active_views;   /* jobs on the device at once (>= 1) */
  int32_t reserved0;
} pm_server_config;

typedef struct pm_job_stats {
  double queue_ms;    /* submit -> admitted */
  double latency_ms;  /* submit -> outputs written */
  int32_t status;     /* the job's pm_status */
  int32_t reserved0;
} pm_job_stats;

typedef struct pm_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t cache_hits, cache_misses, cache_evictions;  /* per source image of a job */
  uint64_t steps, launches;
  uint64_t arena_capacity, arena_used, cache_bytes;    /* bytes */
  int32_t active, pending;
  int32_t max_active_seen;
  int32_t reserved0;
} pm_server_stats;

/* The server reserves hbm_budget_bytes of device memory at creation. */
pm_status pm_server_create(pm_engine* engine, const pm_server_config* config, pm_server** out);
/* Queue one job.  options, view and outputs are copied; the memory they
 * point to (images, cameras, depth maps, output buffers) is borrowed until
 * pm_server_wait returns for the ticket.  Outputs go to the non-null
 * pm_outputs buffers.  PM_ERR_OVER_BUDGET when the job alone needs more
 * than the budget. */
pm_status pm_server_submit(pm_server* server, const pm_options* options, const pm_view_desc* view,
                           const pm_outputs* outputs, uint64_t* ticket);
/* *done = 1 once the job's outputs are written (or it failed). */
pm_status pm_server_poll(pm_server* server, uint64_t ticket, int32_t* done);
/* Block until the job is done; returns its status and releases the ticket. */
pm_status pm_server_wait(pm_server* server, uint64_t ticket, pm_job_stats* stats);
pm_status pm_server_get_stats(pm_server* server, pm_server_stats* out);
/* Finishes every submitted job, then frees the server. */
void pm_server_destroy(pm_server* server);

/* ---- on-disk map formats ----
 * Mat<T>::Write / Read (util/mat.h:178-215): a text header "W&H&D&", then
 * width*height*depth values of elem_size bytes, little-endian, slice-major
 * (DepthMap, conf and cost maps use it as is). */
pm_status pm_mat_write(const char* path, int32_t width, int32_t height, int32_t depth, int32_t elem_size,
                       const void* data);
pm_status pm_mat_read_header(const char* path, int32_t* width, int32_t* height, int32_t* depth);
/* The header must match width, height and depth. */
pm_status pm_mat_read(const char* path, int32_t width, int32_t height, int32_t depth, int32_t elem_size,
                      void* data);
/* NormalMap::Encode / Decode (normal_map.cc:133-176): 3 planar float
 * slices of `pixels` values <-> one uint32 per pixel.  NormalMap::Write is
 * pm_normal_encode + pm_mat_write(..., 1, 4, codes). */
pm_status pm_normal_encode(int64_t pixels, const float* normal, uint32_t* codes);
pm_status pm_normal_decode(int64_t pixels, const uint32_t* codes, float* normal);
/* The problem's normal map encoded on the device (same bits as
 * pm_normal_encode of pm_problem_get_normal_map). */
pm_status pm_problem_get_normal_codes(const pm_problem* problem, uint32_t* out);
/* DepthMap::Write and NormalMap::Write of the problem's current maps, as
 * PatchMatchController::ProcessProblem stores them
 * (patch_match_controller.cc:4706, 4719); either path may be null. */
pm_status pm_problem_write_maps(const pm_problem* problem, const char* depth_path, const char* normal_path);

/* ---- cross-view depth filter ----
 * CrossFilterCuda::Run (patch_match_cuda.cu:1493-1551) with
 * CrossFilterKernel (kernel_functions.cu:3561-3778): every view's depth map
 * is checked against the depth and normal maps of its extended sources
 * (Problem::src_image_extend_idxs).  A pixel keeps a depth when at least
 * min(filter_min_num_consistent, num_src) sources reproject it within
 * filter_geom_consistency_max_cost pixels, with a relative depth difference
 * below depth_diff_threshold and normals within max_normal_error degrees;
 * the kept depth is the reference's median of the consistent depths.  All
 * views are filtered in one device pass. */
typedef struct pm_cross_options {
  float filter_geom_consistency_max_cost;  /* 1.0 (patch_match_options.h:139) */
  float max_normal_error;                  /* degrees, 15 (:141) */
  float depth_diff_threshold;              /* 0.02 (:231) */
  int32_t filter_min_num_consistent;       /* 1 (:135) */
} pm_cross_options;

typedef struct pm_cross_image {
  pm_camera cam;
  int32_t width;
  int32_t height;
  const float* depth;            /* width*height, 0 = no depth */
  const uint32_t* normal_codes;  /* width*height, NormalMap encoding */
} pm_cross_image;

typedef struct pm_cross_view {
  int32_t ref;          /* index into the images */
  int32_t num_src;      /* 1..31 */
  const int32_t* src;   /* indices into the images */
  float* refined;       /* out: the ref image's width*height refined depth map */
} pm_cross_view;

pm_status pm_cross_options_init(pm_cross_options* options);
pm_status pm_cross_filter(pm_engine* engine, const pm_cross_options* options, int32_t num_images,
                          const pm_cross_image* images, int32_t num_views, const pm_cross_view* views);
/* The same filter (CrossFilterCuda::Run, patch_match_cuda.cu:1493-1551) over
 * problems of this engine whose current depth and normal maps are on the
 * device (after pm_problem_run), with each problem's reference camera:
 * views[k].ref and .src index the problems.  The normals are encoded on the
 * device; only the refined maps asked for cross PCIe.  A view whose refined
 * is NULL writes its refined map back into its problem's depth map, after
 * every view has read its sources (two such views may not share a problem). */
pm_status pm_cross_filter_problems(pm_engine* engine, const pm_cross_options* options, int32_t num_problems,
                                   pm_problem* const* problems, int32_t num_views, const pm_cross_view* views);

/* ---- deduplication after the cross filter ----
 * CrossFilterCuda::DedupRun (patch_match_cuda.cu:1561-1612) with FillGpuMat
 * and DeduplicFilterKernel (kernel_functions.cu:3780-3916, 3930-3941), and
 * the getters GetDeduplicDepthMap, GetDeltDepthMap and GetVisibilyMap
 * (patch_match_cuda.cu:1629-1703).  Every image starts with a visibility
 * word of 0x80000000 per pixel.  The views run in the given order, as the
 * reference's problems do: a view's pixel keeps its word (and gains bit
 * 30 - k for each consistent source k) when at least
 * min(filter_min_num_consistent, num_src) sources reproject it within
 * filter_geom_consistency_max_cost pixels and depth_diff_threshold, and no
 * consistent source that is itself a filtered reference (ref_flag) still
 * shows that point as visible; otherwise its word becomes 0.  So a point
 * seen by several filtered views is kept by the last view that sees it.
 * max_normal_error is not used here. */
typedef struct pm_dedup_image {
  pm_camera cam;
  int32_t width;
  int32_t height;
  const float* depth;    /* width*height: the map DedupRun reads (the cross filter's refined map for a
                            filtered reference, the image's own map otherwise; ResetDepthMaps, :1260-1295) */
  int32_t ref_flag;      /* ref_flags_ (CrossFilterCuda ctor): the image is a reference of the cross filter */
  int32_t reserved0;
  uint32_t* visibility;  /* out, may be NULL: width*height visibility words (GetVisibilyMap) */
  float* dedup_depth;    /* out, may be NULL: depth where bit 31 of the word is set, else 0 (GetDeduplicDepthMap) */
} pm_dedup_image;

typedef struct pm_dedup_view {
  int32_t ref;           /* index into the images */
  int32_t num_src;       /* 0..31; a view with no source is skipped (:1586-1588) */
  const int32_t* src;    /* indices into the images, none equal to ref */
  float* delt;           /* out, may be NULL: the ref's width*height GetDeltDepthMap of this view (0 where the
                            pixel is dropped) */
} pm_dedup_view;

pm_status pm_cross_dedup(pm_engine* engine, const pm_cross_options* options, int32_t num_images,
                         const pm_dedup_image* images, int32_t num_views, const pm_dedup_view* views);

/* ---- depth-map fusion ----
 * StereoFusion's fusion of the depth and normal maps into points
 * (fusion.cc:1730-2183, FuseImage and Fuse), over the images in the order
 * StereoFusion::Run takes them (the first non-rig image, then FindNextImage,
 * fusion.cc:697-727, 1150-1215), with one fusing thread per image: seeds
 * row by row at step_size from the 3-pixel margin, each seed walking the
 * consistency graph depth first through the images' overlapping lists.  A
 * pixel belongs to the first walk that reaches it.  The result is the
 * reference's with num_eff_threads = 1 (several threads share the fused
 * pixel masks, so their result depends on timing).  Semantic maps, the ROI
 * box and the steps after the walks (duplicate and outlier removal, plane
 * completion) are not part of this call. */
typedef struct pm_fusion_options {
  int32_t step_size;               /* 1 (fusion.h:57) */
  int32_t min_num_pixels;          /* 5, >= 2 */
  int32_t max_num_pixels;          /* 10000 */
  int32_t min_num_visible_images;  /* 3 */
  int32_t max_traversal_depth;     /* 100 */
  int32_t with_normal;             /* 1 */
  double max_reproj_error;         /* 2.0 px */
  double max_depth_error;          /* 0.01f, relative */
  double max_normal_error;         /* 10.0f degrees */
} pm_fusion_options;

typedef struct pm_fusion_image {
  pm_camera cam;
  int32_t width;
  int32_t height;
  const float* depth;            /* width*height, <= 0 = no depth */
  const float* normal;           /* 3 planar width*height slices in the camera frame; may be NULL without normals */
  const uint8_t* rgb;            /* width*height*channels: interleaved r, g, b (Bitmap::GetPixel), or grey */
  const int32_t* overlapping;    /* overlapping_images_[i]: GetMaxOverlappingImages(check_num_images, 0) */
  int32_t num_overlapping;
  int32_t is_rig;                /* Image::IsRig: FindNextImage takes non-rig images first */
  int32_t used;                  /* used_images_: the image has maps and takes part */
  int32_t channels;              /* 3 (or 0): rgb holds r, g, b; 1: grey, the same value for all three */
} pm_fusion_image;

typedef struct pm_fusion pm_fusion;  /* the fused points of one call */

pm_status pm_fusion_options_init(pm_fusion_options* options);
pm_status pm_fusion_run(pm_engine* engine, const pm_fusion_options* options, int32_t num_images,
                        const pm_fusion_image* images, pm_fusion** out);
/* Number of points and of visibility entries (sum over points). */
pm_status pm_fusion_get_counts(const pm_fusion* fusion, int64_t* num_points, int64_t* num_visibility);
/* The points in the reference's order: xyz (3 per point), normal (3, zero
 * without normals), rgb (3), score (1); visibility as CSR: point p is seen
 * by images vis_images[vis_offsets[p] .. vis_offsets[p+1]) in ascending
 * order.  Any output may be NULL. */
pm_status pm_fusion_get_points(const pm_fusion* fusion, float* xyz, float* normal, uint8_t* rgb, float* score,
                               int64_t* vis_offsets, int32_t* vis_images);
/* How the device got there: walks run, walks that accepted pixels,
 * dataflow launches, walks run alone (their graphs outgrew the device's
 * per-walk buffers), points, wall time of the call; the longest chains of
 * walks that waited one for another, summed over the windows (as the device
 * counted them while it ran: a lower bound, exact unless a depth was read
 * before it landed), and the time the dataflow launches took. */
typedef struct pm_fusion_stats {
  int64_t walks, kept, rounds, big_walks, points;
  double total_ms;
  int64_t chain;
  double flow_ms;
} pm_fusion_stats;
pm_status pm_fusion_get_stats(const pm_fusion* fusion, pm_fusion_stats* out);
void pm_fusion_destroy(pm_fusion* fusion);

/* create + run + read back + destroy in one call. */
pm_status pm_run_view(pm_engine* engine, const pm_options* options,
                      const pm_view_desc* view, const pm_outputs* outputs);

/* ---- sparse model: source-view selection and depth ranges ----
 * The sparse model mvs::Model (mvs/model.h:22-101) as the controller hands
 * it to the patch-match setup (patch_match_controller.cc:3153, 3171, 3294):
 * the registered images' world->camera poses and rig flags, and the sparse
 * points, each with its track (the indices of the images observing it).
 * Tracks are CSR: point p is observed by
 * track_images[track_offsets[p] .. track_offsets[p+1]). */
typedef struct pm_model {
  int32_t num_images;            /* 1..65535 */
  const float* R;                /* num_images x 9, row-major */
  const float* T;                /* num_images x 3 */
  const uint8_t* is_rig;         /* num_images flags (Image::IsRig), NULL = none */
  int64_t num_points;
  const float* xyz;              /* num_points x 3 */
  const int64_t* track_offsets;  /* num_points + 1, non-decreasing, from 0 */
  const int32_t* track_images;   /* track_offsets[num_points] image indices */
} pm_model;

/* Per-phase device times and sizes of one selection (optional output). */
typedef struct pm_overlap_stats {
  int64_t pair_records;   /* (observation, co-observation) records emitted */
  int64_t valid_records;  /* of them, between distinct images of the subset */
  int64_t image_pairs;    /* distinct (image, co-observing image) pairs */
  int32_t fallback_images;/* images ranked by the spatial fallback */
  int32_t reserved0;
  float emit_ms, sort_ms, reduce_ms, rank_ms, total_ms;
  int32_t reserved1;
} pm_overlap_stats;

/* Model::GetMaxOverlappingImages (mvs/model.cc:474-699; with an image
 * subset, :701-934).  For every image of the subset (image_mask[i] != 0;
 * NULL = every image): the co-observing images of the subset whose 75th
 * percentile triangulation angle over the shared points is at least
 * min_triangulation_angle degrees, ranked by the number of shared points
 * (times 1.5, truncated, when both images' rig flags agree), at most
 * max_num_images of them.  An image that observes no point takes the
 * spatial fallback (SelectSpatialNeighborViews): the images whose centre is
 * 0.1..100 mean-nearest-neighbour distances away and whose viewing
 * direction is within 60 degrees, ranked by (1.1 - cos angle) * max(d/m, 1)
 * with near-duplicates of a chosen centre skipped.  Ties rank the lower
 * image index first (the reference leaves their order to its hash map and
 * sort).  Output is CSR: out_offsets[num_images + 1], out_images with room
 * for num_images * max_num_images indices. */
pm_status pm_model_max_overlapping_images(pm_engine* engine, const pm_model* model, int32_t max_num_images,
                                          double min_triangulation_angle, const uint8_t* image_mask,
                                          int32_t* out_offsets, int32_t* out_images, pm_overlap_stats* stats);

/* Model::ComputeDepthRanges (mvs/model.cc:941-960) over the bounding box
 * Model::Read builds (model.cc:221-267: every point and every camera
 * centre): per image the min and max camera z of the box's 8 corners
 * (ComputeDepthRangeFromBBox, model.cc:22-35), min clamped to 1e-6.
 * out: num_images x 2 (depth_min, depth_max). */
pm_status pm_model_depth_ranges(pm_engine* engine, const pm_model* model, float* out);

/* pm_model_depth_ranges without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * All pointers are device memory: d_R num_images x 9 and d_T num_images x 3
 * (world->camera, row-major), d_xyz num_points x 3, d_out num_images x 2
 * (depth_min, depth_max), the same values as pm_model_depth_ranges.  The
 * camera centres and the bounding box live in a fixed device array of the
 * library (one per GPU, held until the process ends), so num_images is at
 * most PM_DEPTH_RANGES_NOALLOC_MAX_IMAGES; more is PM_ERR_INVALID_ARG.  Every
 * call on a GPU shares that array: calls that overlap on one GPU (from
 * several threads or engines) race.  The work runs on the engine's stream,
 * which is synchronised before returning; no hipMalloc, hipFree or
 * allocating library call is made. */
#define PM_DEPTH_RANGES_NOALLOC_MAX_IMAGES 65535
pm_status pm_model_depth_ranges_noalloc(pm_engine* engine, int32_t num_images, const float* d_R, const float* d_T,
                                        int64_t num_points, const float* d_xyz, float* d_out);

#ifdef __cplusplus
}
#endif

#endif  /* SENSEMAP_PM_H_ */
