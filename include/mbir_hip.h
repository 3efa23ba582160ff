/*
 * mbir_hip.h -- thin C-ABI between the C host engine (reconstruct()) and the
 * hand-written gfx950 kernels of the MBIR hot path.
 *
 * Plain pointers and sizes only; no GPU handle crosses this interface: the
 * context owns its device buffers, its HIP stream and its timing events.
 * All host arrays use the reference's layouts:
 *   sinogram-shaped  [N_p][N_r][N_t]  (view, detector column r, row t) --
 *                    the reference's internal order (XT_Init.c:294-301)
 *   object           [N_time][N_z][N_y][N_x]  (no halo slices)
 *   A windows        first[N_y*N_x][N_slots] int32 (-1 = empty column),
 *                    val[N_y*N_x][N_slots][W], W = 2*mult_xy+1, where the
 *                    slots enumerate the views frame by frame
 *                    (slot s holds view frame_views[s]).
 * Every function returns 0 on success and -1 on failure (the reference's
 * convention, XT_Debug.h:22-32); mbir_hip_last_error() gives the reason.
 *
 * Each entry point names the reference function whose role it takes over
 * (paths relative to the reference's Source_Code/ directory).
 */
#ifndef PHASECON_MBIR_HIP_H
#define PHASECON_MBIR_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry and prior of one multi-resolution stage, as set up by
 * initStructures (MBIR_4D/XT_Init.c:211-432). */
typedef struct {
  int32_t N_r, N_t, N_p;           /* detector columns, rows, views       */
  int32_t N_x, N_y, N_z, N_time;   /* voxels and reconstruction frames    */
  int32_t mult_xy, mult_z;         /* voxel scale; z_overlap = mult_z     */
  double delta_r, delta_t, R0, RMax, x0, y0, delta_xy, OffsetR;
  double q, p_s, p_t, c_s, c_t;    /* qGGMRF shape (p split space/time)   */
  double sig_s_q, sig_s_qp, sig_t_q, sig_t_qp; /* sigma^q, sigma^(q-p)    */
  double filt[27], filt_t;         /* neighbour weights [dz][dy][dx]      */
  double alpha;                    /* over-relaxation (XT_Constants.h:87) */
  double huber_T, huber_delta;     /* generalized Huber T, delta          */
  double var_est;                  /* variance parameter sigma^2          */
} mbir_hip_geom;

typedef struct mbir_hip_ctx mbir_hip_ctx;

/* Kernel ids for mbir_hip_kernel_time(). */
enum {
  MBIR_K_AMATRIX = 0,  /* A build                                       */
  MBIR_K_FWDPROJ = 1,  /* A x into the error sinogram                   */
  MBIR_K_SWEEP = 2,    /* parallel ICD sweep                            */
  MBIR_K_COST = 3,     /* FP64 MAP cost                                 */
  MBIR_K_BACKPROJ = 4, /* A^T v                                         */
  MBIR_K_COUNT = 5
};

const char* mbir_hip_last_error(void);

/* Creates a context on the current HIP device and uploads the stage inputs.
 * H is the per-view detector response [N_p][65] (XT_Profile.c:129-218);
 * cosv/sinv the view cosines/sines (XT_Init.c:193-206); frame_ptr
 * [N_time+1] / frame_views [frame_ptr[N_time]] the view->frame map
 * (XT_Init.c:69-96); y, w the projections and weights (w already divided by
 * var_est); d the ring offset [N_r][N_t] or NULL for zero; mask the update
 * mask [N_y][N_x] (XT_ICD_update.c:1127-1136).  N_z is at most 256; the
 * MBIR sweep (mbir_hip_icd_*) takes at most 128 voxels per column and 512
 * views (1024 A values) per column and frame, and refuses more when it
 * runs; the tomographic sweep (mbir_hip_pc_equits) takes the full 256. */
int mbir_hip_create(mbir_hip_ctx** out, const mbir_hip_geom* g,
                    const float* H, const double* cosv, const double* sinv,
                    const int32_t* frame_ptr, const int32_t* frame_views,
                    const float* y, const float* w, const float* d,
                    const uint8_t* mask);
void mbir_hip_destroy(mbir_hip_ctx* ctx);

/* Tuning knobs: "tile" (ICD tile edge in [4, 1024], default 8), "alpha"
 * (override of g->alpha for the sweep) and "afp16" (0 or 1, default 0; takes
 * effect at the next mbir_hip_build_amatrix): A is stored in FP16 with one
 * FP32 scale per voxel column (its largest |a|), and the sweep streams the
 * FP16 values; the FP32 copy of A (mbir_hip_get_amatrix, the back-projector)
 * then holds the same rounded values.  The forward projector and the cost
 * compute A on the fly and stay exact.  The reference stores A in FP32
 * (XT_AMatrix.c:50-110); this knob has no counterpart there.  "nhicd" (0 or
 * 1, default 0): every sweep records each visited column's sum of |dx| in a
 * magnitude map (MagUpdateMap, XT_VoxUpdate.c:158,265), which
 * mbir_hip_nhicd ranks.  "placement" (default 0): how a sweep launch's
 * (tile, frame) work items map to workgroups -- 0 tile-major, 1 each frame
 * bound to one XCD (the hardware deals workgroups to the 8 XCDs
 * round-robin; with 8 frames per launch every frame's error-sinogram
 * traffic then stays in one L2; with other frame counts a frame spans
 * 8 / gcd(frames, 8) XCDs; measured slower on the 4D benchmark).
 * "xcd_check" (0 or 1): the sweep counts its work items and those that ran
 * on another XCD than workgroup id mod 8 (mbir_hip_xcd_check).
 * "pair_parities" (0 or 1, default 0; placement 0): an equit's 8 (colour,
 * frame parity) launches become 4, launch k running colour k on the even
 * frames together with colour k + 1 on the odd ones -- sets with no
 * neighbouring voxels and disjoint views, so one launch holds twice the
 * work and the visiting order stays a sequential one.  With 4x4 tiles on
 * small grids the paired sweep needs a lower over-relaxation than the
 * colour-major one (measured: 64x64x4 x 4 frames diverges at alpha >= 1.8
 * paired, converges at 1.9 unpaired).  None of these has a counterpart in
 * the reference. */
int mbir_hip_set_option(mbir_hip_ctx* ctx, const char* key, double value);

/* Counters of option "xcd_check" since it was set: sweep work items run and
 * those that ran on another XCD than their placement intended. */
int mbir_hip_xcd_check(mbir_hip_ctx* ctx, int64_t* items, int64_t* off_xcd);

/* Object transfer (replaces the Object array of XT_Structures.h:215). */
int mbir_hip_set_object(mbir_hip_ctx* ctx, const float* x);
int mbir_hip_get_object(mbir_hip_ctx* ctx, float* x);

/* The object of the coarse stage up-sampled on the device into the fine
 * stage (the next multi-resolution stage: twice N_x and N_y, the same or
 * twice N_z, the same frames), by the reference's rule for the next stage's
 * initial object (XT_ICD_update.c:358-390 in x-y, :448-496 in x-y-z: ends
 * copied, interior samples 3/4 and 1/4 of the two nearest).  masked = 1 also
 * zeroes the voxels outside the fine stage's update mask, which the
 * reference leaves at their interpolated values and never updates; with it
 * the fine stage's ICD solves the same problem as a start from zero. */
int mbir_hip_upsample_object(mbir_hip_ctx* fine, mbir_hip_ctx* coarse,
                             int32_t masked);

/* Saves (restore = 0) or restores (restore = 1) the object on the device,
 * for the cost-rise guard of the equit loops (mbir_guard_check in
 * mbir_engine.h); no reference counterpart (the reference stops at a cost
 * rise, XT_ICD_update.c:1171).  The error sinogram is not saved: recompute
 * it with mbir_hip_init_error after a restore. */
int mbir_hip_object_snapshot(mbir_hip_ctx* ctx, int32_t restore);

/* Builds the stored system matrix for every (column, view) pair;
 * replaces the per-visit calcAMatrixColumnforAngle (XT_AMatrix.c:50-110).
 * get_amatrix returns it in the window layout above (values dequantised
 * when stored in FP16). */
int mbir_hip_build_amatrix(mbir_hip_ctx* ctx);
int mbir_hip_get_amatrix(mbir_hip_ctx* ctx, int32_t* first, float* val);

/* e = y - A x - d from the current object, A without the delta_t factor
 * (initErrorSinogam, XT_ICD_update.c:651-698), accumulated in FP64 and
 * rounded once.  Calling it again between sweeps removes the rounding drift
 * the sweeps' FP32 atomic updates accumulate in e. */
int mbir_hip_init_error(mbir_hip_ctx* ctx);
int mbir_hip_get_error(mbir_hip_ctx* ctx, float* e);
int mbir_hip_set_error(mbir_hip_ctx* ctx, const float* e);

/* Pure projector pair for tests and the warm start:
 * ax = A x  (forward_project_voxel_AMat1D, XT_ForwardProject.c:44-61) and
 * bt = A^T v over each frame's views (the adjoint of that scatter). */
int mbir_hip_forward_project(mbir_hip_ctx* ctx, const float* x, float* ax);
int mbir_hip_backproject(mbir_hip_ctx* ctx, const float* v, float* bt);

/* Normalised back-projection, the input of the learned warm start
 * (include/unet_hip.h; no counterpart in the reference):
 * out[t][z][y][x] = A^T(w (y - d)) / A^T(w A m) at the voxel, m the update
 * mask's indicator, over its frame's views (0 where the denominator is 0):
 * exact for an object that is uniform inside the mask.  Needs A
 * (mbir_hip_build_amatrix); the projection A m is computed on the fly. */
int mbir_hip_bp_normalized(mbir_hip_ctx* ctx, float* out);

/* Filtered back-projection, the learned warm start's default input (no
 * counterpart in the reference): out[t][z][y][x] = A^T(h * (y - d)) / K
 * over each frame's views, h the Ram-Lak filter along the detector channels
 * (h(0) = 1/4, h(n) = -1/(pi n)^2 for odd n, 0 for even n != 0) and K the
 * mean over the update mask of A^T(h * A m), so that an object uniform
 * inside the mask reconstructs to its value.  Needs A. */
int mbir_hip_fbp(mbir_hip_ctx* ctx, float* out);

/* n equits of the parallel ICD (updateVoxelsTimeSlices,
 * XT_ICD_update.c:932-1092, with the voxel update of XT_VoxUpdate.c:45-283
 * and the substitution of XT_ICD_update.c:133-182).  Iteration numbers run
 * first_iter .. first_iter+n-1 (zero skipping starts at 2).  stats (may be
 * NULL) receives [sum |dx|, sum |x_new|, zero-skipped voxels, updated
 * voxels] over the n equits. */
int mbir_hip_icd_equits(mbir_hip_ctx* ctx, int32_t n, int32_t first_iter,
                        uint32_t seed, double* stats);

/* The same n equits captured once into a HIP graph (one persistent launch
 * sequence: the 8 sweep launches of every equit, graph-safe clears, no memset
 * nodes) and replayed on each call; the iteration number comes from device
 * memory, so one graph serves every first_iter.  The graph is captured
 * again when n, seed, the tile, the A storage or alpha change.  Replays are
 * asynchronous (mbir_hip_synchronize waits).  No counterpart in the
 * reference, which runs the same schedule as host loops
 * (XT_ICD_update.c:932-1092). */
int mbir_hip_icd_graph(mbir_hip_ctx* ctx, int32_t n, int32_t first_iter, uint32_t seed);

/* Non-homogeneous ICD passes (the NHICD loops of updateVoxelsTimeSlices,
 * XT_ICD_update.c:1007-1046, with the voxel selection of
 * VSC_based_Voxel_Line_Select, XT_NHICD.c:164-216).  Each pass filters the
 * magnitude map with the 5x5 Hamming window (ComputeVSC, XT_NHICD.c:126-160)
 * and then runs the sweep's 8 launches with every workgroup visiting only
 * the floor(fraction n) (at least one) of its tile's n columns whose
 * filtered magnitude is highest; the visited columns' magnitudes are
 * refreshed for the next pass.  The reference selects the top fraction of
 * each frame's lines with a quickselect threshold; the per-tile selection
 * here keeps the workgroups of a launch equally long.  Needs option "nhicd"
 * = 1 and a tile of at most 16 x 16 columns.  iter is the iteration number
 * (zero skipping from 2); stats as for mbir_hip_icd_equits, over the
 * passes. */
int mbir_hip_nhicd(mbir_hip_ctx* ctx, int32_t passes, double fraction, int32_t iter,
                   uint32_t seed, double* stats);

/* Projection step of the phase-contrast phantom simulator (ForwardProject,
 * PCMBIR_3D/XT_genSinogram.c:146-186, on the GPU): the absorption (mag)
 * and refractive-index (phase) phantoms [n_z][n_xy][n_xy] projected at the
 * phantom's own voxel pitch, detector N_r = n_xy channels by N_t = n_z rows
 * (the geometry of initPhantomStructures, PCMBIR_3D/XT_Init.c:611-690, with
 * n_xy / n_z in place of PHANTOM_XY_SIZE / PHANTOM_Z_SIZE): proj_real /
 * proj_imag [n_p][n_z][n_xy] = sum over voxels of phantom x A x wavenumber
 * x the voxel line response along z (PCMBIR_3D/XT_Profile.c:182-290).
 * n_xy must be a multiple of proj_cols. */
int mbir_hip_phantom_project(const float* mag, const float* phase, int32_t n_xy, int32_t n_z,
                             int32_t n_p, const float* angles, int32_t proj_rows,
                             int32_t proj_cols, double vox_wid, double rot_center,
                             double wavenumber, float* proj_real, float* proj_imag);

/* Down-sampling of already normalised data, the standard-format driver's
 * read_ProjWeightData (reconstruct/real_data/Standard_Data_Format/
 * XT_HDFIO.c:90-111), on the GPU from the caller's HDF5 hyperslabs:
 * proj / weights_in [n_proj][rows ratio][cols ratio] float.  Per output
 * sample, t = the mean of its ratio x ratio block of projections and
 * projections = bh_coef t^2 + t (BH_QUAD_COEF = 0 for this driver),
 * weights = the block's sum of weights; outputs [n_proj][rows][cols], in
 * the reference's float / double arithmetic. */
int mbir_hip_std_downsample(const float* proj, const float* weights_in, int32_t n_proj,
                            int32_t rows, int32_t cols, int32_t ratio, double bh_coef,
                            float* projections, float* weights);

/* Flat-field normalisation of raw APS counts (read_data,
 * reconstruct/real_data/APS_Data_Format/XT_HDFIO.c:138-178), on the GPU,
 * from host buffers the caller has read from its HDF5 files (the
 * reference's hyperslabs: white and dark frames 1 .. N-2, the rows from
 * datafile_row0, the channels centred): proj [n_proj][rows ratio][cols
 * ratio], white [n_white][..], dark [n_dark][..] uint16.  The white and
 * dark frames are averaged per pixel; per output sample (view, row,
 * channel), over its ratio x ratio block of pixels, weights = sum |proj -
 * dark|, and with t = log(sum |white - dark| / weights), projections =
 * bh_coef t^2 + t (BH_QUAD_COEF = 0.5 in the reference).  Outputs
 * [n_proj][rows][cols] float, in the reference's arithmetic (float sums in
 * its order, double log and polynomial). */
int mbir_hip_aps_normalize(const uint16_t* proj, const uint16_t* white, const uint16_t* dark,
                           int32_t n_proj, int32_t n_white, int32_t n_dark, int32_t rows,
                           int32_t cols, int32_t ratio, double bh_coef, float* projections,
                           float* weights);

/* The magnitude map [N_time][N_y][N_x] (filtered = 0) or its VSC filtering
 * (filtered = 1).  Needs option "nhicd" = 1. */
int mbir_hip_get_magnitudes(mbir_hip_ctx* ctx, int32_t filtered, float* out);

/* One ring-offset update (update_Sinogram_Offset, XT_ICD_update.c:915-923)
 * of the offset d [N_r][N_t] and the error sinogram together (e + d is
 * unchanged): mode 1 unconstrained (update_d_offset_unconstrained,
 * :866-911), 2 with sum d = 0 (:813-863), 3 with <A_m, d> = 0 for the
 * rectangular-patch windows A_m of XT_OffsetError.c:46-96
 * (:749-810 with constrained_quad_opt, XT_OffsetError.c:99-165).  Entries
 * count with the generalized-Huber weight of their current residual
 * (w where |e sqrt(w)| < huber_T, huber_T huber_delta sqrt(w) / |e|
 * otherwise); the reference uses the selection stored at its last update.
 * d starts as the create() argument (zero when NULL) and is also used by the
 * forward projector from then on. */
int mbir_hip_offset_update(mbir_hip_ctx* ctx, int32_t mode);
int mbir_hip_get_offset(mbir_hip_ctx* ctx, float* d);

/* The noise-variance re-estimate (update_variance_parameter,
 * XT_ICD_update.c:716-746): var_new = mean over the sinogram of e^2 W
 * (selected entries) or |e| huber_delta huber_T sqrt(W var) (the others),
 * W = w var the measured weight; w becomes W / var_new and g->var_est
 * var_new (the cost's log term).  *var_est (may be NULL) receives it. */
int mbir_hip_variance_update(mbir_hip_ctx* ctx, double* var_est);
int mbir_hip_get_weight(mbir_hip_ctx* ctx, float* w);

/* FP64 MAP cost (computeCost, XT_ICD_update.c:185-318).  fresh = 1
 * recomputes e = y - A x - d in FP64 on the fly; fresh = 0 uses the error
 * sinogram the sweeps maintain; fresh = 2 is an estimate for stopping tests:
 * the maintained e and each prior pair's potential in FP32 (within about
 * 1e-6 relative), summed in FP64.  out = [total, data term incl.
 * N log(var)/2, prior term]. */
int mbir_hip_cost(mbir_hip_ctx* ctx, int32_t fresh, double* out);

/* Batched functional substitution (CE_FunctionalSubstitution,
 * XT_ICD_update.c:133-182) on n independent voxels, FP32: nb [n][27]
 * ([dz][dy][dx], centre ignored), nbf its flags, tn/tf [n][2] the t-1/t+1
 * neighbours and flags.  No context needed. */
int mbir_hip_substitution_batch(const mbir_hip_geom* g, int32_t n,
                                const float* V, const float* theta1,
                                const float* theta2, const float* nb,
                                const uint8_t* nbf, const float* tn,
                                const uint8_t* tf, float* out);

/* mbir_hip_substitution_batch without any allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc); no counterpart in the reference.
 * The inputs and the output are staged through one fixed device array per
 * GPU (8 MiB, held by the library from module load to process end), so n is
 * at most MBIR_SUBST_NOALLOC_MAX (else -1 with a message).  Calls that
 * overlap on one GPU (several host threads) share that array and race;
 * sequential calls are independent (every staged byte is rewritten). */
#define MBIR_SUBST_NOALLOC_MAX 49152
int mbir_hip_substitution_batch_noalloc(const mbir_hip_geom* g, int32_t n,
                                        const float* V, const float* theta1,
                                        const float* theta2, const float* nb,
                                        const uint8_t* nbf, const float* tn,
                                        const uint8_t* tf, float* out);

/* theta1/theta2 of n voxels (t, z, y, x quadruples) from the stored A, the
 * current error sinogram and weights, FP32 accumulation
 * (XT_VoxUpdate.c:56-85). */
int mbir_hip_voxel_thetas(mbir_hip_ctx* ctx, int32_t n, const int32_t* tzyx,
                          float* theta1, float* theta2);

/* ---- PCMBIR tomographic ICD (the ENABLE_TOMO_RECONS path of PCMBIR_3D,
 * compiled out of the shipped reference).  A context switched to this mode
 * holds two objects, delta (PhaseObject) and beta (MagObject), and their
 * error sinograms e = target - k A x (A the MBIR rule of the context's
 * geometry, k the wavenumber: PCMBIR_3D/XT_AMatrix.c:103), and runs the
 * tomographic ICD: theta1 = -mu sum e k a, theta2 = mu sum (k a)^2
 * (PCMBIR_3D/XT_VoxUpdate.c:23-80), the coupled substitution of both
 * objects (FunctionalSubstitution, PCMBIR_3D/XT_Prior.c:137-239), zero
 * skipping when both are zero.  Index 0 of the pairs is the phase (delta)
 * prior, 1 the magnitude (beta) prior; T is DecorrTran row-major
 * (PCMBIR_3D/XT_Init.c:234-246); sigmas already raised to q and q - p and
 * scaled as PCMBIR_3D/XT_Init.c:482-489 does; filt / filt_t as initFilter
 * (XT_Init.c:144-176: 1/distance normalised to 1, temporal 0). */
typedef struct {
  double mu, wavenumber, alpha;
  double T[4];
  double c_s[2], sig_s_q[2], sig_s_qp[2], c_t[2], sig_t_q[2], sig_t_qp[2];
  double p, q;
  double filt[27], filt_t;
} mbir_hip_pc_prior;

int mbir_hip_pc_enable(mbir_hip_ctx* ctx, const mbir_hip_pc_prior* prior);
/* Objects [N_time][N_z][N_y][N_x]. */
int mbir_hip_pc_set_objects(mbir_hip_ctx* ctx, const float* delta, const float* beta);
int mbir_hip_pc_get_objects(mbir_hip_ctx* ctx, float* delta, float* beta);
/* e = target - k A x of both objects (internal layout [N_p][N_r][N_t]),
 * FP64 accumulation rounded once. */
int mbir_hip_pc_init_error(mbir_hip_ctx* ctx, const float* target_delta,
                           const float* target_beta);
int mbir_hip_pc_get_error(mbir_hip_ctx* ctx, float* e_delta, float* e_beta);
/* n equits (iterations first_iter ..); stats (may be NULL) receives
 * [sum |dx| over both objects, voxels updated, voxels skipped]. */
int mbir_hip_pc_equits(mbir_hip_ctx* ctx, int32_t n, int32_t first_iter, uint32_t seed,
                       double* stats);
/* ADMM head loop of the tomographic mode (PCMBIR_3D/XT_ICD_update.c:1037-1151
 * with ENABLE_TOMO_RECONS).  The context keeps, per object, the phase
 * retrieval's projection estimate Z, the scaled dual V and the previous
 * estimate P (FP32, internal layout [N_p][N_r][N_t] = the phase retrieval's
 * [n_views][rows][cols]); the error sinograms then hold e = Z - V - k A x.
 * init (initErrorSinogam, XT_ICD_update.c:726-760): P = z, Z = z or, with
 * from_object (recon_type 2), Z = k A x of the current objects, V = 0.
 * target: a = Z - e = k A x + V, the retrieval's proximal point
 * (XT_ICD_update.c:1089-1093).  update with the retrieval's new estimate z
 * (XT_ICD_update.c:1095-1140): V' = V + k A x - z, e = z - V' - k A x, P = Z
 * = z; residuals (may be NULL) = [beta primal, beta dual, delta primal,
 * delta dual], primal = RMS of z - k A x, dual = |z - P| / |V'|. */
int mbir_hip_pc_admm_init(mbir_hip_ctx* ctx, const double* z_delta, const double* z_beta,
                          int32_t from_object);
int mbir_hip_pc_admm_target(mbir_hip_ctx* ctx, double* a_delta, double* a_beta);
int mbir_hip_pc_admm_update(mbir_hip_ctx* ctx, const double* z_delta, const double* z_beta,
                            double* residuals);
/* which: 0 Z, 1 V, 2 P. */
int mbir_hip_pc_admm_state(mbir_hip_ctx* ctx, int32_t which, float* delta, float* beta);
/* FP64 cost: out = [total, mu/2 sum(e_delta^2 + e_beta^2), prior] (the
 * reference's computeCost squares mu e, PCMBIR_3D/XT_ICD_update.c:92-102;
 * the same at mu = 1). */
int mbir_hip_pc_cost(mbir_hip_ctx* ctx, double* out);

/* Device time of the kernels of one id since the last reset, from HIP
 * events recorded on the context's stream around every launch. */
int mbir_hip_kernel_time(mbir_hip_ctx* ctx, int32_t kernel_id, double* ms,
                         int64_t* launches);
int mbir_hip_reset_kernel_time(mbir_hip_ctx* ctx);
int mbir_hip_synchronize(mbir_hip_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
