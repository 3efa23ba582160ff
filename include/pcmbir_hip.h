/*
 * pcmbir_hip.h -- C-ABI of the gfx950 kernels of the PCMBIR phase-retrieval
 * path, over all views at once: the Fresnel transform and its Hermitian, the
 * steepest-descent phase retrieval of w, the Nelder-Mead pass and the ADMM
 * projection estimate (estimate_complex_projection), and the single-material
 * Paganin retrieval.  phcontomo_reconstruct (include/pcmbir4d.h) drives it.
 *
 * Paths of reference files are relative to the reference's
 * Source_Code/PCMBIR_3D/ directory.  Complex arrays are interleaved float
 * pairs (re, im) laid out [n_views][rows][cols], rows = N_r, cols = N_t, the
 * reference's per-view [N_r][N_t] order.  rows and cols are powers of two in
 * [2, 1024] (the reference's FFTW transform is circular on N_r x N_t, Q14;
 * callers that want padding pad before the call).  Every function returns 0
 * on success and -1 on failure, with the reason in pcmbir_hip_last_error().
 */
#ifndef PHASECON_PCMBIR_HIP_H
#define PHASECON_PCMBIR_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcmbir_hip_ctx pcmbir_hip_ctx;

const char* pcmbir_hip_last_error(void);

/* Context for n_views transforms of rows x cols with sample spacings
 * delta_rows / delta_cols, wavelength lam and propagation distance dist (same
 * length unit).  window [rows][cols] is the frequency window of H
 * (create_FresnelTranWindow, XT_Init.c:210-231), NULL for none; H^H applies no
 * window, as the reference's does not (Q11). */
int pcmbir_hip_create(pcmbir_hip_ctx** out, int32_t n_views, int32_t rows,
                      int32_t cols, double delta_rows, double delta_cols,
                      double lam, double dist, const float* window);
void pcmbir_hip_destroy(pcmbir_hip_ctx* ctx);

/* y = H x (herm = 0, compute_FresnelTran, XT_FresnelTran.c:12-78) or
 * y = H^H x (herm = 1, compute_HermFresnelTran, :82-137) for every view; x
 * and y may be the same array. */
int pcmbir_hip_fresnel(pcmbir_hip_ctx* ctx, int32_t herm, const float* x,
                       float* y);

/* Options: "fft_reg" 1 (default) runs the transforms' passes with each
 * sequence in registers (radix-8 Stockham, LDS only between passes) where a
 * sequence has at least 8 values, 0 the in-LDS radix-2 passes.
 * "fft_col_threads" (a power of two, 64-1024, default 512): the most
 * threads of one workgroup of the register-resident column pass, rows / 8
 * per column (512: 8 columns of 512 rows). */
int pcmbir_hip_set_option(pcmbir_hip_ctx* ctx, const char* key, double value);

/* Phase-retrieval problem of every view (XT_PhaseRet.c): measurements y, the
 * factor D, weights lambda [n_views][rows][cols] (real), proximal target v,
 * initial estimate w and phase Omega, penalty nu.  omega = NULL sets Omega to
 * the phase projection of w (compute_phase_projection, XT_PhaseRet.c:10-33),
 * as initErrorSinogam does before the phase retrieval
 * (PCMBIR_3D/XT_ICD_update.c:788-789). */
int pcmbir_hip_set_phase_problem(pcmbir_hip_ctx* ctx, const float* y,
                                 const float* D, const float* lambda,
                                 const float* v, const float* w,
                                 const float* omega, double nu);

/* Steepest descent on w, every view in parallel (steepest_descent_iter,
 * XT_PhaseRet.c:207-270, run by the loop of XT_CmplxProjEst.c:165-169): a
 * view stops after an iteration j > 1 whose update 100 sum|alpha g| / sum|w|
 * is below thresh, or after max_iters.  iters [n_views] receives the
 * iterations each view ran, upd [n_views] its last update (either may be
 * NULL). */
int pcmbir_hip_sd_run(pcmbir_hip_ctx* ctx, int32_t max_iters, double thresh,
                      int32_t* iters, double* upd);

/* Current w and Omega of every view. */
int pcmbir_hip_get_phase_state(pcmbir_hip_ctx* ctx, float* w, float* omega);

/* Cost |y - Omega H D w|^2_lambda / 2 + nu |w - v|^2 / 2 of every view
 * (compute_cost, XT_PhaseRet.c:172-205), FP64 accumulation. */
int pcmbir_hip_phase_cost(pcmbir_hip_ctx* ctx, double* cost);

/* Projection-domain prior of the simplex search (Cost_NMS without
 * ENABLE_TOMO_RECONS, XT_Search.c:8-43 -- the reference never defines it,
 * Q12): phase and magnitude qGGMRF sigma and c (XT_Init.c:482-490), the
 * decorrelating transform [[1, d/b], [1, -d/b]] when decorrelate (recon_type
 * 2, XT_Init.c:234-246) else the identity, qGGMRF p (MRF_P = 1.2, q = 2),
 * and the ADMM nu.  The search uses the reference's constants rho 1, chi 2,
 * gamma 0.5, sigma 0.5, threshold 0.005 %, 50 iterations (XT_Init.c:507-519). */
int pcmbir_hip_set_nms_prior(pcmbir_hip_ctx* ctx, double phase_sig,
                             double phase_c, double mag_sig, double mag_c,
                             double delta_over_beta, int32_t decorrelate,
                             double p, double nu);

/* The tomographic coupling of the phase retrieval (the ENABLE_TOMO_RECONS
 * branches of Cost_NMS, XT_Search.c:18-19, and compute_pretcost,
 * XT_CmplxProjEst.c:38-41, compiled out of the shipped build, Q12): with a
 * target set, the simplex search and the pretcost use mu/2 |z - a|^2 with a
 * = (a_mag, a_phase) [n_views][rows][cols] FP64 (the ADMM head loop's
 * projection plus dual, PCMBIR_3D/XT_ICD_update.c:1089-1093) in place of the
 * projection-domain prior.  a_mag or a_phase NULL switches the coupling off. */
int pcmbir_hip_set_tomo_target(pcmbir_hip_ctx* ctx, double mu, const double* a_mag,
                               const double* a_phase);

/* One Nelder-Mead pass (Nelder_Mead_Simplex_2DSearch, XT_Search.c:100-215)
 * over every pixel of every view with b = w - v given: zr, zi
 * [n_views][rows][cols] FP64 in and out, b complex.  Pixels are visited in
 * four colours (row parity, column parity), each colour's pixels in
 * parallel; the reference visits them in raster order.  iters_total (may be
 * NULL) receives the summed search iterations. */
int pcmbir_hip_nms_pass(pcmbir_hip_ctx* ctx, double* zr, double* zi,
                        const float* b, int64_t* iters_total);

/* The phase-retrieval ADMM of every view (estimate_complex_projection,
 * XT_CmplxProjEst.c:105-229) on the problem of pcmbir_hip_set_phase_problem
 * (its v is replaced by the loop's steepest-descent target): v = 0, then up
 * to pret_iters of { NMS pass on z with b = w - v; steepest descent on w
 * (sd_iters, sd_thresh); v += exp(-z) - w; residuals; pretcost }, a view
 * stopping after an iteration > 1 whose dual residual is below pret_thresh.
 * zr, zi in and out; per view iters, cost_init / cost_final (pretcost,
 * XT_CmplxProjEst.c:15-103) and the histories of the residuals primal / dual
 * and of the pretcost after each iteration, costs, all [n_views][pret_iters]
 * with the first iters[v] entries of view v written (any output may be
 * NULL). */
int pcmbir_hip_estimate(pcmbir_hip_ctx* ctx, double* zr, double* zi,
                        int32_t pret_iters, double pret_thresh,
                        int32_t sd_iters, double sd_thresh, int32_t* iters,
                        double* cost_init, double* cost_final, double* primal,
                        double* dual, double* costs);

/* The ADMM dual v of every view after pcmbir_hip_estimate (the reference's
 * MagPRetDual / PhasePRetDual), complex [n_views][rows][cols]; zeros before
 * the first estimate. */
int pcmbir_hip_get_dual(pcmbir_hip_ctx* ctx, float* v);

/* Single-material Paganin phase retrieval of every view
 * (paganins_1mat_phase_retrieval, PCMBIR_3D/XT_Paganin.c:77-134): intensity
 * |Re y|^2 / |D|^2, filtered by 1 / (1 + z (delta/beta) / (2 k) 4 pi^2
 * (u^2 + v^2)) with the context's distance z and spacings, thickness
 * -log|Re|; zr = thickness / 2, zi = thickness (delta/beta) / 2, FP64
 * [n_views][rows][cols].  y, D complex; k = the wavenumber 2 pi / lambda. */
int pcmbir_hip_paganin(pcmbir_hip_ctx* ctx, const float* y, const float* D,
                       double delta_over_beta, double wavenumber, double* zr,
                       double* zi);

/* Device time (ms) and launch count of the Fresnel FFT passes since the
 * last call, from HIP events on the context's stream. */
int pcmbir_hip_fft_time(pcmbir_hip_ctx* ctx, double* ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
