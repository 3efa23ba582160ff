/*
 * mbir4d.h -- drop-in replacement for the reference's MBIR 4D library header
 * (Source_Code/include/mbir4d.h:35).
 *
 * reconstruct() keeps the reference's argument list, argument meaning,
 * return codes (0 / -1), CWD output files and their byte layouts, and the
 * returned object layout [recon_num][proj_rows+2][proj_cols][proj_cols]
 * float with the reconstruction in z slices 1..proj_rows and zero halo
 * slices 0 and proj_rows+1 (MBIR_4D/XT_Init.c:362-363,
 * MBIR_4D/XT_Engine.c:170).  The reference header declares it void; its
 * definition (MBIR_4D/XT_Engine.c:81) returns int, as does this one.
 * The caller owns every input and frees *object with free().
 *
 * Differences from the reference, by design of this library:
 *   - single process, one GPU: no MPI; proj_rows is not split across ranks;
 *   - the ICD voxel order differs (parallel ICD on the GPU), so results are
 *     compared on the converged MAP cost, not per-iteration equality;
 *   - remove_rings 1-3 (ring offset estimation) and, with quad_convex = 0,
 *     the noise-variance estimate are updated once per equit from the
 *     second stage on (the reference: the offset after every half-sweep of
 *     its NHICD subsets);
 *   - a cost rise is handled as the reference does (-1 and "Cost value
 *     increased.", MBIR_4D/XT_ICD_update.c:1171) only when it exceeds 1e-6
 *     relative (the FP32 rounding noise of the maintained error sinogram),
 *     is confirmed from a recomputed error sinogram, and persists after two
 *     recoveries; each recovery restores the object of the lowest cost in
 *     the stage and halves the over-relaxation's excess over 1, with a WARN
 *     line in the log.
 */
#ifndef MBIR4D_H_
#define MBIR4D_H_
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

int reconstruct(float** object, float* projections, float* weights,
                float* proj_angles, float* proj_times, float* recon_times,
                int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                int32_t recon_num, float vox_wid, float rot_center,
                float sig_s, float sig_t, float c_s, float c_t,
                float convg_thresh, int32_t remove_rings, int32_t quad_convex,
                float huber_delta, float huber_T, uint8_t restart,
                FILE* debug_msg_ptr);

#ifdef __cplusplus
}
#endif
#endif
