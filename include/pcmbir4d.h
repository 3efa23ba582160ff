/*
 * pcmbir4d.h -- drop-in replacement for the reference's PCMBIR library
 * header (Source_Code/include/pcmbir4d.h:35-36).
 *
 * phcontomo_reconstruct() keeps the reference's argument list, argument
 * meaning, return codes (0 / -1) and CWD output files with their byte
 * layouts (PCMBIR_3D/XT_Engine.c:82-196, PCMBIR_3D/XT_ICD_update.c:956-1212).
 * As in the shipped reference, whose tomographic reconstruction is compiled
 * out (ENABLE_TOMO_RECONS is never defined), it retrieves projections and
 * returns no object: *magobject and *phaseobject are set to NULL.
 *   recon_type 1: single-material Paganin retrieval of every view
 *                 (pag_mag_ret.bin, pag_phase_ret.bin);
 *   recon_type 2: the per-view ADMM phase retrieval
 *                 (estimate_complex_projection) started from the Paganin
 *                 estimate read back from those files (mag_tomo_aux.bin,
 *                 phase_tomo_aux.bin and the auxiliary files).
 * The caller owns every input.
 *
 * Differences from the reference, by design of this library:
 *   - single process, one GPU: no MPI; proj_rows is not split across ranks;
 *   - proj_rows and proj_cols must be powers of two up to 1024 (the GPU
 *     transforms are radix-2; the reference's FFTW takes any size);
 *   - all views run together and the simplex search visits pixels in four
 *     colours instead of raster order;
 *   - recon_type 2 without readable pag_*_ret.bin files returns -1;
 *   - no TIFF files are written.
 *
 * phcontomo_forward_project() (PCMBIR_3D/XT_Engine.c:199-247, the phantom
 * simulator of XT_genSinogram.c:59-343) keeps the reference's prototype and
 * reads the two 128 x 512 x 512 float phantoms from the files named by the
 * environment variables PHCONTOMO_MAG_PHANTOM and PHCONTOMO_PHASE_PHANTOM
 * (the reference's are the fixed paths MAG_PHANTOM_FILEPATH /
 * PHASE_PHANTOM_FILEPATH, PCMBIR_3D/XT_Constants.h:91-92).  It writes
 * SimMagProjs.bin, SimPhaseProjs.bin, mag_phantom.bin and phase_phantom.bin
 * as the reference does; the measurement noise has the reference's
 * distribution (m + sqrt(m) N(0, 1), absolute value) from another
 * generator, so noisy measurements match in distribution, not bit for bit.
 */
#ifndef PCMBIR4D_H_
#define PCMBIR4D_H_
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

int phcontomo_reconstruct(float** magobject, float** phaseobject,
                          float* measurements, float* brights,
                          float* proj_angles, float* proj_times,
                          float* recon_times, int32_t proj_rows,
                          int32_t proj_cols, int32_t proj_num,
                          int32_t recon_num, float vox_wid, float rot_center,
                          float mag_sig_s, float mag_sig_t, float mag_c_s,
                          float mag_c_t, float phase_sig_s, float phase_sig_t,
                          float phase_c_s, float phase_c_t, float convg_thresh,
                          float obj2det_dist, float light_energy,
                          float pag_regparam, uint8_t restart,
                          uint8_t recon_type, FILE* debug_msg_ptr);

int phcontomo_forward_project(float** measurements, float** brights,
                              float* proj_angles, int32_t proj_rows,
                              int32_t proj_cols, int32_t proj_num,
                              float vox_wid, float rot_center,
                              float obj2det_dist, float light_energy,
                              float pag_regparam, FILE* debug_msg_ptr);

/* The simulator on caller-supplied phantoms mag / phase [n_z][n_xy][n_xy]
 * (n_xy, n_z powers of two up to 1024 and multiples of proj_cols /
 * proj_rows; the reference's are 512 and 128): measurements
 * [proj_num][proj_rows][proj_cols] and brights [proj_rows][proj_cols];
 * noise_seed 0 leaves the noise out.  No counterpart in the reference,
 * whose simulator reads fixed files. */
int phcontomo_forward_project_ex(float* measurements, float* brights, const float* mag,
                                 const float* phase, int32_t n_xy, int32_t n_z,
                                 const float* proj_angles, int32_t proj_rows,
                                 int32_t proj_cols, int32_t proj_num, float vox_wid,
                                 float rot_center, float obj2det_dist,
                                 float light_energy, uint64_t noise_seed, FILE* dbg);

#ifdef __cplusplus
}
#endif
#endif
