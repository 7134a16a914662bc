/* pm_oracle.h -- CPU transcription of SenseMap's ACMM PatchMatch stage.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker that the gfx950 HIP
 * engine is compared against, and the CPU baseline that bench.py times; the
 * product never links or calls it.
 *
 * It restates, in plain C with the canonical FP32 evaluation order shared
 * with the HIP kernels, the semantics of:
 *   src/mvs/gpu_mat_ref_image.cu:16-56      FilterKernel (ref moments)
 *   src/mvs/kernel_functions.cu:70-306      random draws, perturbation, CorrectNormal
 *   src/mvs/kernel_functions.cu:409-568     PropagateDepth, ComposeHomography,
 *                                           ComputeGeomConsistencyCostShared
 *   src/mvs/kernel_functions.cu:636-686     FindMinCost, TransformPDFToCDF
 *   src/mvs/kernel_functions.cu:689-737     InitNormalMap, InitDepthMap
 *   src/mvs/kernel_functions.cu:1245-1340   ComputeNCCCost<R,S>
 *   src/mvs/kernel_functions.cu:1447-1601   ComputeMultiViewCostVector, InitConfMapKernel
 *   src/mvs/kernel_functions.cu:1603-1651   ComputeCurvatureMap (+ BubbleSort, utils.h:206-221)
 *   src/mvs/kernel_functions.cu:1655-1675   ComputeGradientMap
 *   src/mvs/kernel_functions.cu:2136-2560   SweepFromCheckerBoardOpt
 *   src/mvs/kernel_functions.cu:2736-2931   LocalOptimization
 *   src/mvs/kernel_functions.cu:3080-3093   MaskDepthMapKernel
 *   src/mvs/kernel_functions.cu:3095-3280   ComputeGeomGradientMap, MedianFilter,
 *                                           Black/RedPixelFilter
 *   src/mvs/kernel_functions.cu:3282-3354   FilterDepthMap
 *   src/mvs/patch_match_cuda.cu:493-580     InitTransforms (pose packing)
 *   src/mvs/patch_match_cuda.cu:582-698     InitWorkspace
 *   src/mvs/patch_match_cuda.cu:748-1057    Run() iteration schedule
 *   src/mvs/utils.cc:14-52                  relative pose / projection matrices
 *
 * plus an FP64 restatement of the NCC cost (pmo_ncc_cost_f64) used as the
 * precision reference.
 *
 * Pinned independently (tests/test_oracle_*.py): the random generator
 * (Random123 known answers), the math functions (libm FP64), the pose
 * packing (closed forms), the reference moments (FP64 brute force), the NCC
 * cost (closed forms: identical, affine and inverted windows, flat windows,
 * integer shifts), the geometric term, the sweep's candidate gathering
 * (pmo_gather_candidates: a designed conf map whose picks follow from the
 * reference's offset table), the view selection and weighted argmin
 * (pmo_select_views: hand-computed cases), the refinement hypotheses and
 * LocalOptimization's trials (pmo_refine_hypotheses, pmo_lo_trials: the
 * ranges they are drawn from, over many pixels), the filters and the
 * sparse-model selection (pm_model_oracle.c).
 * LocalOptimization's choice among its trials and the schedule's threshold
 * (pmo_lo_decide, pmo_th_mc) are pinned by hand-computed cases.  The
 * curvature map (pmo_curvature_map) is pinned by closed forms (constant and
 * ramp images, a normal field rotating per column), scipy's correlation, a
 * float64 brute-force median and a designed case of NaN angles; the mask
 * (pmo_mask_depth) by its skipped last row and column and its place after
 * FilterDepthMap in a run.
 * LocalOptimization's cost refresh for the sources it did not select is
 * pinned in photometric runs: every stored cost equals a fresh FP32 NCC at
 * the final maps (a transcription without the refresh fails that).  Its
 * geometric term, which reproduces the reference passing ref_invK as K
 * (kernel_functions.cu:2924), is pinned by the closed form that K gives:
 * the back-projection lands near the image origin, so away from it the
 * term saturates at geom_consistency_max_cost and every stored cost is
 * either NCC + lambda * (true-K term) or NCC + lambda * max_cost.  The
 * order of the phases (pmo_schedule) is pinned against the reference's
 * order written down in the tests, composed phase by phase with
 * pmo_run_phases; red-first, early or reordered local-optimisation
 * schedules give other maps there.
 */
#ifndef SENSEMAP_PM_ORACLE_H_
#define SENSEMAP_PM_ORACLE_H_

#include <stdint.h>
#include "../include/sensemap_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-view precomputed data, filled by pmo_prepare. */
typedef struct pmo_view {
  int W, H, N;
  const uint8_t* ref;          /* W*H, tightly packed copy */
  const uint8_t* src[PM_MAX_NUM_SRC];
  int srcW[PM_MAX_NUM_SRC], srcH[PM_MAX_NUM_SRC];
  const float* src_depth[PM_MAX_NUM_SRC];
  float refK[4];               /* fx, cx, fy, cy */
  float refInvK[4];            /* 1/fx, -cx/fx, 1/fy, -cy/fy */
  float pose[PM_MAX_NUM_SRC][PM_POSE_STRIDE];
  float depth_min, depth_max;
  uint32_t view_id;
  /* bilateral weights: spatial table (per window sample) and colour table */
  int nwin;                    /* samples per window row */
  float* sp;                   /* nwin*nwin */
  float ct[256];
  /* per-pixel ref moments */
  float* mu_r;
  float* var_r;
  float* inv_sw;
  float* err_r;                /* sum w (r - mu) / sum w, compensated mode */
  uint8_t* grad_ok;            /* ComputeGradientMap >= 1.0 */
  const pm_view_desc* desc;    /* the view as given (for the FP64 cost restatement) */
} pmo_view;

typedef struct pmo_state {
  float* depth;                /* W*H */
  float* normal;               /* 3*W*H planar */
  float* conf;                 /* W*H */
  float* cost;                 /* N*W*H */
  uint32_t* view_sel;          /* W*H */
} pmo_state;

typedef struct pmo_counters {
  uint64_t ncc_evaluations;
  uint64_t ncc_samples;
  uint64_t geo_evaluations;
} pmo_counters;

/* Cost precision of the runs that follow (process-wide; for the precision
 * study): 0, the default, is the FP32 transcription; 1 computes every NCC
 * cost with the FP64 restatement (pmo_ncc_cost_f64) and rounds it to FP32
 * for the decisions, everything else unchanged: the "FP64 run" the FP32
 * runs' depth maps are compared with. */
void pmo_set_cost_fp64(int on);

/* Full Run() of one view (photometric or geometric per options).  Pixels in
 * rows [row_begin, row_end) are updated; the others keep their initial state
 * (row_end <= 0 means all rows).  num_threads <= 0 uses the OpenMP default.
 * Returns 0 on success. */
int pmo_run_view(const pm_options* opt, const pm_view_desc* view, const pm_outputs* out,
                 int row_begin, int row_end, int num_threads, pmo_counters* counters);

/* Phases of Run(), for checking its order: kind (init, sweep, local
 * optimisation), iteration and colour (0 = black = (row + col) even). */
#define PMO_PHASE_INIT 0
#define PMO_PHASE_SWEEP 1
#define PMO_PHASE_LOCAL 2
#define PMO_PHASE(kind, iter, color) ((int32_t)(((kind) << 16) | ((iter) << 1) | (color)))
#define PMO_PHASE_KIND(p) ((int)((p) >> 16))
#define PMO_PHASE_ITER(p) ((int)(((p) & 0xffff) >> 1))
#define PMO_PHASE_COLOR(p) ((int)((p) & 1))
/* The phase list pmo_run_view runs (before the steps after the
 * iterations); returns its length, or -1 when capacity is short. */
int pmo_schedule(const pm_options* opt, int32_t* phases, int capacity);
/* Runs exactly the given phases on the whole view, from the state
 * InitWorkspace sets up (zeros, or the given init maps), with no steps after
 * them.  Returns 0 on success. */
int pmo_run_phases(const pm_options* opt, const pm_view_desc* view, const pm_outputs* out, const int32_t* phases,
                   int num_phases, int num_threads);
/* Geometric term (reference K) of every pixel's depth against every source:
 * N planar W*H, NaN where depth <= 0. */
int pmo_geom_cost_maps(const pm_options* opt, const pm_view_desc* view, const float* depth, int num_threads,
                       float* out);

/* Pose packing exactly as the engine does it (see pm_problem_get_poses). */
int pmo_pack_poses(const pm_view_desc* view, float* poses, float ref_k8[8]);

/* Ref moments for a whole image: mu, var, 1/sum_w, gradient gate. */
int pmo_ref_moments(const pm_options* opt, const uint8_t* img, int W, int H,
                    float* mu, float* var, float* inv_sw, uint8_t* grad_ok);

/* One NCC evaluation (FP32 canonical) and its FP64 restatement, for the
 * hypothesis {nx, ny, nz, depth} at (row, col) against source `src`. */
float pmo_ncc_cost(const pm_options* opt, const pm_view_desc* view, int row, int col,
                   const float hyp[4], int src);
double pmo_ncc_cost_f64(const pm_options* opt, const pm_view_desc* view, int row, int col,
                        const double hyp[4], int src);
/* Both costs for every pixel and source of the hypothesis maps (depth
 * W*H, normal 3 planar W*H) into out32 / out64 (N planar W*H, NaN where
 * depth <= 0). */
int pmo_cost_maps(const pm_options* opt, const pm_view_desc* view, const float* depth, const float* normal,
                  int num_threads, float* out32, double* out64);
/* RunFilter on view->init_depth into depth_out (W*H). */
int pmo_run_filter(const pm_options* opt, const pm_view_desc* view, int dedup, float* depth_out, int num_threads);
/* CrossFilterCuda::Run over the views (refined maps into views[k].refined). */
/* StereoFusion's walk (pm_fusion_oracle.c), same options and images as the
 * C-ABI's pm_fusion_run; NULL on failure. */
struct pmo_fusion;
struct pmo_fusion* pmo_fusion_run(const pm_fusion_options* options, int num_images, const pm_fusion_image* images);
int pmo_fusion_counts(const struct pmo_fusion* f, int64_t* num_points, int64_t* num_vis);
int pmo_fusion_points(const struct pmo_fusion* f, float* xyz, float* normal, uint8_t* rgb, float* score,
                      int64_t* vis_offsets, int32_t* vis_images);
void pmo_fusion_free(struct pmo_fusion* f);
/* CrossFilterCuda::DedupRun and its getters over the views in order, same
 * arguments and outputs as the C-ABI's pm_cross_dedup. */
int pmo_cross_dedup(const pm_cross_options* options, int num_images, const pm_dedup_image* images, int num_views,
                    const pm_dedup_view* views, int num_threads);
int pmo_cross_filter(const pm_cross_options* options, int num_images, const pm_cross_image* images,
                     int num_views, const pm_cross_view* views, int num_threads);
/* Iteration iter's good-matching threshold, and LocalOptimization's choice
 * among its 5 trials given their cost rows (see pm_oracle.c). */
float pmo_th_mc(const pm_options* opt, int iter);
int pmo_lo_decide(const float* M, uint32_t sel, int N, uint32_t live, float conf, float* conf_out);
/* Refinement hypotheses of sweep pixel (row, col) after its num_samples
 * view-selection draws, and LocalOptimization's 5 trials of a pixel (see
 * pm_oracle.c): out gets 5 rows {nx, ny, nz, d}. */
int pmo_refine_hypotheses(const pm_options* opt, uint32_t view_id, uint32_t pass, int W, int row, int col,
                          const float* iK, float depth, const float* normal, float depth_min, float depth_max,
                          float* out);
int pmo_lo_trials(int W, int H, const float* depth, const float* normal, int row, int col, uint64_t seed,
                  uint32_t view_id, uint32_t pass, float* out);
/* Candidate hypotheses of one sweep pixel (see pm_oracle.c): out gets up to
 * 11 rows {nx, ny, nz, d}; returns the count. */
int pmo_gather_candidates(int W, int H, const float* conf, const float* depth, const float* normal, int grad_ok,
                          const float* iK, int row, int col, float* out);
/* View selection and weighted argmin of one sweep pixel (see pm_oracle.c):
 * M is nh rows of PM_MAX_NUM_SRC costs, draws num_samples uniforms.  Returns
 * the chosen hypothesis, or -1 when the pixel keeps its state. */
int pmo_select_views(const pm_options* opt, float th_mc, int nh, int N, const float* M, const float* draws,
                     float* wts, uint32_t* sel);
/* Model::GetMaxOverlappingImages and Model::ComputeDepthRanges
 * (pm_model_oracle.c), same arguments and outputs as the C-ABI's
 * pm_model_max_overlapping_images / pm_model_depth_ranges. */
int pmo_model_max_overlapping_images(const pm_model* m, int max_num_images, double min_triangulation_angle,
                                     const uint8_t* mask, int32_t* out_offsets, int32_t* out_images);
int pmo_model_depth_ranges(const pm_model* m, float* out);
/* Geometric-consistency term (FP32 canonical), K = ref {fx, cx, fy, cy}. */
float pmo_geom_cost(const pm_options* opt, const pm_view_desc* view, int row, int col,
                    float depth, int src);

/* Deterministic math and RNG, exported for the tests. */
/* ComputeCurvatureMap (kernel_functions.cu:1603-1651): out is 2 x W x H
 * planar {gradient, curvature}; ref is the W x H 8-bit reference image and
 * normal the 3 x W x H planar normal map. */
int pmo_curvature_map(const uint8_t* ref, int W, int H, const float* normal, uint16_t* out);
/* Run()'s median filter (patch_match_cuda.cu:991-1008: ComputeGeomGradientMap,
 * BlackPixelFilter, RedPixelFilter; kernel_functions.cu:3095-3280) in place
 * on a W x H depth map and its 3 x W x H planar normal map. */
int pmo_median_filter(float* depth, float* normal, int W, int H, float diff_thres);
/* MaskDepthMapKernel (kernel_functions.cu:3080-3093) on a W x H depth map. */
int pmo_mask_depth(const uint8_t* mask, int W, int H, float* depth);
float pmo_expf(float x);
float pmo_sinf(float x);
float pmo_cosf(float x);
float pmo_atan2f(float y, float x);
float pmo_acosf(float x);
void pmo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float pmo_uniform(uint64_t seed, uint32_t view, uint32_t pixel, uint32_t pass, uint32_t draw);

#ifdef __cplusplus
}
#endif

#endif
