/*
 * mbir_engine.h -- host-side stage set-up of the MBIR engine, exported so
 * callers (tests, the benchmark, other drivers) can drive the HIP C-ABI of
 * mbir_hip.h one stage at a time, and reconstruct_ex(), the extended form of
 * reconstruct() (mbir4d.h).
 *
 * Each helper restates one piece of the reference's initStructures
 * (MBIR_4D/XT_Init.c:211-432) for a single process; paths are relative to
 * the reference's Source_Code/ directory.
 */
#ifndef PHASECON_MBIR_ENGINE_H
#define PHASECON_MBIR_ENGINE_H
#include <stdint.h>
#include <stdio.h>

#include "mbir_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MBIR_MAX_STAGES 8 /* MAX_MULTRES_NUM, MBIR_4D/XT_Constants.h:100 */

/* Multi-resolution voxel scales, coarse to fine (MBIR_4D/XT_Engine.c:
 * 111-141).  Returns the number of stages. */
int32_t mbir_multires_schedule(int32_t proj_rows, int32_t proj_cols,
                               int32_t* mult_xy, int32_t* mult_z);

/* Time-interlaced view schedule of the APS driver
 * (gen_interlaced_views_0_to_Inf,
 * reconstruct/real_data/APS_Data_Format/XT_Views.c:7-46).  Candidate view i
 * has index i K + bitrev(floor(i / L) mod K) (L = N_theta / K, the bit
 * reversal over round(log2 K) bits), angle index * pi / N_theta and time
 * index / K; a candidate is kept when its angle exceeds the last kept
 * one's by more than rot_speed * min_acq_time (the first always is).
 * Kept views proj_start .. proj_start + N_p - 1 go to views / times.
 * Returns proj_start + N_p, or -1 on a bad argument. */
int32_t mbir_interlaced_views(float* views, float* times, int32_t K, int32_t N_theta,
                              int32_t proj_start, int32_t N_p, float min_acq_time,
                              float rot_speed);

/* View -> frame map (MBIR_4D/XT_Init.c:69-96).  frame_ptr[N_time+1],
 * frame_views[<= N_p].  Returns -1 if times decrease (XT_Init.c:109-113),
 * else the number of assigned views. */
int32_t mbir_view_map(int32_t N_p, const float* proj_times, int32_t N_time,
                      const float* recon_times, int32_t* frame_ptr,
                      int32_t* frame_views);

/* Geometry and prior of one stage (MBIR_4D/XT_Init.c:138-190, 211-432);
 * p_s/p_t are the spatial/temporal qGGMRF p (the reference has one,
 * MRF_P = 1.2).  Returns -1 on the reference's divisibility errors. */
int mbir_stage_geometry(mbir_hip_geom* g, int32_t proj_rows,
                        int32_t proj_cols, int32_t proj_num, int32_t recon_num,
                        float vox_wid, float rot_center, float sig_s,
                        float sig_t, float c_s, float c_t, int32_t mult_xy,
                        int32_t mult_z, int32_t quad_convex, float huber_delta,
                        float huber_T, const int32_t* frame_ptr, double p_s,
                        double p_t);

/* Per-view detector response H[N_p][65] (MBIR_4D/XT_Profile.c:47-218). */
void mbir_detector_response(const mbir_hip_geom* g, const float* angles,
                            float* H);

/* View cosines and sines (MBIR_4D/XT_Init.c:193-206). */
void mbir_sincos(int32_t n, const float* angles, double* cosv, double* sinv);

/* Inscribed-circle update mask [N_y][N_x] of radius vox_wid*proj_cols/2
 * (MBIR_4D/XT_Init.c:243, XT_ICD_update.c:1127-1136). */
void mbir_update_mask(const mbir_hip_geom* g, double radius, uint8_t* mask);

/* File layout [p][t][r] -> internal [p][r][t] (MBIR_4D/XT_Init.c:294-301). */
void mbir_sino_to_internal(int32_t N_p, int32_t N_t, int32_t N_r,
                           const float* in, float* out);

/* Options of reconstruct_ex(); mbir_default_options() gives the values
 * reconstruct() uses. */
typedef struct {
  double p_s, p_t;         /* qGGMRF p, spatial / temporal (1.2, 1.2)     */
  int32_t tile;            /* ICD tile edge (8)                           */
  double alpha;            /* over-relaxation, <= 0: reference's 1.5      */
  int32_t max_equits;      /* per stage (1000, MAX_NUM_ITERATIONS)        */
  int32_t single_stage;    /* 1: only the full-resolution stage           */
  int32_t write_files;     /* 1: write the reference's CWD output files   */
  double stop_cost;        /* > 0: also stop once the FP64 cost <= this   */
  uint32_t seed;           /* ICD visiting-order seed (761521)            */
  int32_t refresh_every;   /* exact e = y - A x recompute period (10)     */
} mbir_options;

void mbir_default_options(mbir_options* opt);

/* Cost-rise guard of an equit loop.  The reference stops with -1 at any
 * rise of its FP64 cost (check_error, MBIR_4D/XT_ICD_update.c:1171); the
 * parallel sweep's cost is usually evaluated from the error sinogram its
 * FP32 atomics maintain, which carries rounding noise, and its asynchronous
 * updates can overshoot when over-relaxed close to 2.  The guard accepts a
 * cost within a relative tol of the lowest accepted one, asks for a cost
 * from a recomputed error sinogram before it calls a larger one a rise, and
 * on a confirmed rise asks the caller to restore the object saved at the
 * lowest cost and halve alpha - 1, at most max_recoveries times; after that
 * it fails (the reference's -1). */
typedef struct {
  double best;             /* lowest accepted cost (that of the saved object) */
  double tol;              /* relative rise taken as rounding noise           */
  int32_t recoveries;      /* recoveries made                                 */
  int32_t max_recoveries;
} mbir_guard;

enum {
  MBIR_GUARD_OK = 0,       /* accepted                                        */
  MBIR_GUARD_SAVE = 1,     /* accepted, lowest so far: save the object        */
  MBIR_GUARD_CONFIRM = 2,  /* call again with the cost of a recomputed e      */
  MBIR_GUARD_RECOVER = 3,  /* restore the saved object, recompute e, lower alpha */
  MBIR_GUARD_FAIL = 4      /* rise after max_recoveries: stop with an error   */
};

void mbir_guard_init(mbir_guard* gd, double cost0, double tol, int32_t max_recoveries);
/* exact = 0: cost from the maintained error sinogram; 1: from a recomputed
 * one.  Returns one of the MBIR_GUARD_ codes. */
int32_t mbir_guard_check(mbir_guard* gd, double cost, int32_t exact);
/* The over-relaxation after a recovery: 1 + (alpha - 1) / 2 (alpha > 1), else
 * alpha / 2. */
double mbir_guard_alpha(double alpha);

/* reconstruct() with options.  equits_out / cost_out (may be NULL) receive
 * the equits run in the last stage and the final FP64 cost. */
int reconstruct_ex(float** object, float* projections, float* weights,
                   float* proj_angles, float* proj_times, float* recon_times,
                   int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                   int32_t recon_num, float vox_wid, float rot_center,
                   float sig_s, float sig_t, float c_s, float c_t,
                   float convg_thresh, int32_t remove_rings,
                   int32_t quad_convex, float huber_delta, float huber_T,
                   uint8_t restart, FILE* debug_msg_ptr,
                   const mbir_options* opt, int32_t* equits_out,
                   double* cost_out);

#ifdef __cplusplus
}
#endif
#endif
