/*
 * mbir_oracle.h -- CPU restatement of the reference MBIR_4D hot path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the HIP path is tested against:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (phasecon_amd/libmbir4d.so) never links or
 * calls it.
 *
 * Every function names the reference file:line it restates (paths relative
 * to the reference's Source_Code/ directory).  How each one is pinned (golden
 * vectors from the reference build in oracle/_ref, closed forms, invariants)
 * is listed in the header of oracle/mbir_oracle.c.
 *
 * Array layouts used throughout:
 *   sinogram-shaped (y, w, e):  [N_p][N_r][N_t]       (view, column r, row t)
 *   object (x):                 [N_time][N_z][N_y][N_x]   (no halo slices)
 *   A windows:                  first[col][view] (int32, -1 = empty),
 *                               val[col][view][W] (W = 2*mult_xy+1)
 *                               with col = y*N_x + x.
 */
#ifndef PHASECON_MBIR_ORACLE_H
#define PHASECON_MBIR_ORACLE_H
#include <stdint.h>

#define ORC_RESPONSE_BINS 64   /* XT_Constants.h:79 */

typedef struct {
  /* detector / sinogram */
  int32_t N_r, N_t, N_p;
  double delta_r, delta_t, R0, RMax, T0, TMax, OffsetR, OffsetT;
  /* object */
  int32_t N_x, N_y, N_z, N_time, mult_xy, mult_z, W;
  double x0, y0, z0, delta_xy, delta_z, radius_obj, delta_recon;
  /* prior (separate spatial / temporal p; the reference has one MRF_P) */
  double p_s, p_t, q, c_s, c_t;
  double sig_s_q, sig_s_qp, sig_t_q, sig_t_qp;
  double filt[27];  /* spatial filter [dz+1][dy+1][dx+1], centre 0 */
  double filt_t;    /* temporal filter */
  double alpha;     /* over-relaxation */
  /* forward model */
  double huber_T, huber_delta, var_est;
} orc_geom;

#ifdef __cplusplus
extern "C" {
#endif

/* Park-Miller generator (randlib.c:119-141, 168-175). */
void orc_srandom2(uint32_t seed);
double orc_random2(void);

/* View -> frame map (XT_Init.c:69-96).  frame_ptr has N_time+1 entries,
 * frame_views has up to N_p entries.  Returns the total assigned views. */
int32_t orc_view_map(int32_t N_p, const float* proj_times, int32_t N_time,
                     const float* recon_times, int32_t* frame_ptr,
                     int32_t* frame_views);

/* Geometry and prior set-up for one multi-resolution stage
 * (XT_Init.c:138-190, 211-432).  Returns 0, or -1 on the reference's
 * divisibility errors. */
int32_t orc_setup(orc_geom* g, int32_t proj_rows, int32_t proj_cols,
                  int32_t proj_num, int32_t recon_num, float vox_wid,
                  float rot_center, float sig_s, float sig_t, float c_s,
                  float c_t, int32_t mult_xy, int32_t mult_z,
                  int32_t quad_convex, float huber_delta, float huber_T,
                  const int32_t* frame_ptr, double p_s, double p_t);

/* Per-view detector response table H[N_p][65] (XT_Profile.c:47-124,
 * 129-218), float like the reference's Real_arr_t table. */
void orc_detector_response(const orc_geom* g, const float* angles, float* H);

/* cos/sin of the view angles (XT_Init.c:193-206). */
void orc_sincos(int32_t N_p, const float* angles, double* cosv, double* sinv);

/* One A column (XT_AMatrix.c:50-110).  Returns the count; idx/val hold the
 * kept bins in increasing order. */
int32_t orc_amatrix_column(const orc_geom* g, const double* cosv,
                           const double* sinv, const float* H, int32_t row,
                           int32_t col, int32_t view, int32_t* idx,
                           double* val);

/* Projection of the PCMBIR simulator (ForwardProject,
 * PCMBIR_3D/XT_genSinogram.c:146-186) in the geometry of
 * initPhantomStructures (PCMBIR_3D/XT_Init.c:611-690) with a phantom of
 * n_z x n_xy x n_xy voxels ([z][y][x]) in place of the 128 x 512 x 512 of
 * PHANTOM_Z_SIZE / PHANTOM_XY_SIZE: out_r / out_i [n_p][n_z][n_xy] =
 * sum over voxels of phantom x A(r) k x L(z, t), A the detector response
 * column (XT_AMatrix.c, times the wavenumber k) and L the voxel line
 * response (ZLineResponseProfile / storeVoxelLineResponse,
 * PCMBIR_3D/XT_Profile.c:182-290), accumulated in FP64.  Returns 0, or -1
 * if n_xy is not a multiple of proj_cols. */
int32_t orc_phantom_project(const float* mag, const float* phase, int32_t n_xy,
                            int32_t n_z, int32_t n_p, const float* angles,
                            int32_t proj_rows, int32_t proj_cols, double vox_wid,
                            double rot_center, double wavenumber, double* out_r,
                            double* out_i);

/* All columns for all views in window form. */
void orc_amatrix_all(const orc_geom* g, const double* cosv, const double* sinv,
                     const float* H, int32_t* first, double* val);

/* Forward projection (XT_ForwardProject.c:44-61 via XT_ICD_update.c:651-682)
 * ax[p][r][t] = sum_voxels x * A (no delta_t factor). */
void orc_forward_project(const orc_geom* g, const double* cosv,
                         const double* sinv, const float* H,
                         const int32_t* frame_ptr, const int32_t* frame_views,
                         const double* x, double* ax);

/* qGGMRF derivative-over-delta b(delta) with the delta=0 limit
 * (XT_ICD_update.c:92-128); spatial=1 selects the spatial parameters. */
double orc_qggmrf_weight(const orc_geom* g, double delta, int spatial);

/* qGGMRF potential rho(delta) (XT_ICD_update.c:82-90). */
double orc_qggmrf_value(const orc_geom* g, double delta, int spatial);

/* Functional substitution (XT_ICD_update.c:133-182).  nb[27] is the
 * spatial neighbourhood [dz][dy][dx], nbf[27] its in-bounds flags (the
 * centre is ignored), tn[2] = x at t-1, t+1, tf[2] their flags. */
double orc_substitution(const orc_geom* g, double V, double theta1,
                        double theta2, const double* nb, const uint8_t* nbf,
                        const double* tn, const uint8_t* tf);

/* theta1/theta2 of one voxel over its sparse column
 * (XT_VoxUpdate.c:56-85), with the generalized-Huber reweighting. */
void orc_voxel_thetas(const orc_geom* g, const double* cosv,
                      const double* sinv, const float* H,
                      const int32_t* frame_ptr, const int32_t* frame_views,
                      const double* e, const double* w, int32_t t, int32_t z,
                      int32_t yy, int32_t xx, double* theta1, double* theta2);

/* MAP cost in FP64 (XT_ICD_update.c:185-318).  out[0] = data term incl.
 * N*log(var)/2, out[1] = prior term; returns their sum. */
double orc_cost(const orc_geom* g, const double* e, const double* w,
                const double* x, double* out);

/* Inscribed-circle update mask [N_y][N_x] (XT_ICD_update.c:1127-1136). */
void orc_mask(const orc_geom* g, uint8_t* mask);

/* One ICD equit with the reference's task decomposition
 * (XT_ICD_update.c:499-531, 932-1047; XT_VoxUpdate.c:119-283):
 * random line order per (frame, z-block) from orc_random2, K line subsets,
 * two checkerboard phases over (block parity xor frame parity).  NHICD is
 * not restated (use_nhicd must be 0).  x and e are updated in place.
 * stats[0] = sum |dx|, stats[1] = sum |x_new|, stats[2] = zero-skipped.
 * Deterministic for any OpenMP thread count. */
void orc_icd_equit(const orc_geom* g, const double* cosv, const double* sinv,
                   const float* H, const int32_t* frame_ptr,
                   const int32_t* frame_views, const uint8_t* mask,
                   int32_t num_z_blocks, int32_t K, int32_t iter, double* x,
                   double* e, const double* w, double* stats);

/* Lock-step group simulation of a concurrent ICD schedule (test-only
 * study tool, no reference counterpart): every step updates one column per
 * active tile (tiles of T x T columns, 2x2 tile checkerboard, frame parity)
 * from the same error-sinogram snapshot, then applies all updates.  z is
 * processed in two parity halves inside the column as the GPU kernel does.
 * alpha overrides g->alpha.  Returns nothing; x and e updated in place. */
void orc_icd_equit_tiled(const orc_geom* g, const double* cosv,
                         const double* sinv, const float* H,
                         const int32_t* frame_ptr, const int32_t* frame_views,
                         const uint8_t* mask, int32_t T, double alpha,
                         uint32_t seed, int32_t iter, double* x, double* e,
                         const double* w);

/* As orc_icd_equit_tiled, but each concurrent wavefront works on a private
 * copy of the error sinogram for m consecutive columns of its tile and the
 * copies' updates are merged every m columns (m = 1 is the lock-step
 * model above).  Small problems only: one sinogram copy per thread. */
void orc_icd_equit_tiled_sync(const orc_geom* g, const double* cosv,
                              const double* sinv, const float* H,
                              const int32_t* frame_ptr,
                              const int32_t* frame_views, const uint8_t* mask,
                              int32_t T, double alpha, uint32_t seed,
                              int32_t iter, int32_t m, double* x, double* e,
                              const double* w);

/* Model of the GPU sweep (test-only study tool, no reference counterpart):
 * the kernel's launch order, tiles, work-item placement on cache groups and
 * column visiting order, all items of a launch in lock-step, and a column's
 * theta computed without the other items' updates of the last lag_same
 * steps (same group) or lag_cross steps (other groups; < 0: the whole
 * launch so far).  lag_same = lag_cross = 0 with one item per frame per
 * launch is the sequential ICD in the kernel's order.  stats as
 * orc_icd_equit. */
typedef struct {
  int32_t T;           /* tile edge */
  double alpha;        /* over-relaxation */
  uint32_t seed;       /* the sweep call's seed */
  int32_t iter;        /* iteration number (zero skipping from 2) */
  int32_t call_index;  /* equit index within the sweep call */
  int32_t lag_same, lag_cross;
  int32_t groups;      /* cache groups (XCDs) */
  int32_t placement;   /* 0: tile-major grid, group = id mod groups;
                          1: frames bound to groups */
} orc_async_opts;

void orc_icd_equit_async(const orc_geom* g, const double* cosv,
                         const double* sinv, const float* H,
                         const int32_t* frame_ptr, const int32_t* frame_views,
                         const uint8_t* mask, const orc_async_opts* op,
                         double* x, double* e, const double* w,
                         double* stats);

/* PCMBIR tomographic ICD (row a14; PCMBIR_3D code compiled out of the
 * shipped reference).  Index 0 of the paired arrays is the phase (delta)
 * prior, 1 the magnitude (beta) prior; T is DecorrTran row-major. */
typedef struct {
  double mu, wavenumber;
  double T[4];
  double c_s[2], sq_s[2], sqp_s[2], c_t[2], sq_t[2], sqp_t[2];
  double p, q;
  double filt[27], filt_t;
  double alpha;
} orc_pc_prior;

/* FunctionalSubstitution (PCMBIR_3D/XT_Prior.c:137-239): updates *Vd, *Vb. */
void orc_pc_substitution(const orc_pc_prior* P, double* Vd, double* Vb, double th1d,
                         double th1b, double th2d, double th2b, const double* nbd,
                         const double* nbb, const uint8_t* nbf, const double* tnd,
                         const double* tnb, const uint8_t* tf);
/* mu/2 sum(e_d^2 + e_b^2) + the decorrelated prior; out = [data, prior]. */
double orc_pc_cost(const orc_geom* g, const orc_pc_prior* P, const double* ed,
                   const double* eb, const double* xd, const double* xb, double* out);
/* One equit of the GPU sweep's order, serial (PCMBIR_3D/XT_VoxUpdate.c). */
void orc_pc_equit(const orc_geom* g, const double* cosv, const double* sinv, const float* H,
                  const int32_t* frame_ptr, const int32_t* frame_views, const uint8_t* mask,
                  const orc_pc_prior* P, int32_t T, uint32_t seed, int32_t iter,
                  int32_t call_index, double* xd, double* xb, double* ed, double* eb,
                  double* stats);

/* File-layout <-> internal sinogram transpose (XT_Init.c:294-301):
 * file [p][t][r] -> internal [p][r][t]. */
void orc_sino_from_file(int32_t N_p, int32_t N_t, int32_t N_r,
                        const float* in, double* out);

/* Ring-offset windows of the rectangular-patch constraint
 * (gen_offset_constraint_windows, XT_OffsetError.c:46-96, one node): if win
 * is not NULL it receives [M][N_r][N_t] weights.  Returns M. */
int32_t orc_offset_windows(int32_t N_r, int32_t N_t, double* win);

/* One ring-offset update (update_Sinogram_Offset, XT_ICD_update.c:915-923)
 * of mode 1 (update_d_offset_unconstrained, :866-911), 2
 * (update_d_offset_zero_mean_constraint, :813-863) or 3
 * (update_d_offset_rect_patch_constraint, :749-810, with
 * constrained_quad_opt, XT_OffsetError.c:99-165).  e [N_p][N_r][N_t] and
 * d [N_r][N_t] are updated in place, sel (the generalized-Huber selection,
 * 1 where |e sqrt(w)| < T) is used and then recomputed from the new e. */
void orc_offset_update(int32_t mode, int32_t N_p, int32_t N_r, int32_t N_t,
                       double T, double delta, const double* w, double* e,
                       double* d, uint8_t* sel);

/* The noise-variance re-estimate (update_variance_parameter,
 * XT_ICD_update.c:716-746) over n sinogram entries: w (already divided by
 * *var) is scaled to w * var_old / var_new, sel used then recomputed, *var
 * replaced by the new estimate. */
void orc_variance_update(int64_t n, double T, double delta, double* w,
                         const double* e, uint8_t* sel, double* var);

#ifdef __cplusplus
}
#endif
#endif
