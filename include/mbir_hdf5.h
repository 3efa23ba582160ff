/*
 * mbir_hdf5.h -- host-side HDF5 input stages of the reference's real-data
 * drivers, for one process (libmbir4d_hdf5.so; needs libhdf5).
 *
 * Each reads the same hyperslabs as the driver's own reader and hands them
 * to the GPU normalisation of libmbir4d (mbir_hip_aps_normalize,
 * mbir_hip_std_downsample).  Outputs are the driver's arrays for
 * reconstruct(): projections and weights [proj_num][proj_rows][proj_cols]
 * float (channel fastest).  Return 0, or -1 with a message on
 * debug_file_ptr.  Paths of reference files are relative to its
 * Source_Code/ directory.
 */
#ifndef PHASECON_MBIR_HDF5_H
#define PHASECON_MBIR_HDF5_H
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The APS driver's read_data
 * (reconstruct/real_data/APS_Data_Format/XT_HDFIO.c:11-190), same
 * arguments: uint16 counts /exchange/data, /exchange/data_white and
 * /exchange/data_dark ([frames][rows][channels]); white and dark frames 1
 * .. N-2 averaged; rows datafile_row0 .. + proj_rows * ratio; the channels'
 * remainder modulo proj_cols split evenly off both sides; ratio = channels /
 * proj_cols; beam-hardening coefficient 0.5 (BH_QUAD_COEF,
 * APS_Data_Format/XT_Constants.h:7).  Unlike the reference it also refuses
 * proj_start + proj_num beyond the data's frames. */
int32_t mbir_read_aps_hdf5(const char* data_filename, const char* whites_filename,
                           const char* darks_filename, float* projections,
                           float* weights, int32_t datafile_row0, int32_t proj_rows,
                           int32_t proj_cols, int32_t proj_start, int32_t proj_num,
                           FILE* debug_file_ptr);

/* The standard driver's read_data
 * (reconstruct/real_data/Standard_Data_Format/XT_Main.c:88-93):
 * read_ProjWeightData (XT_HDFIO.c:11-130; float /projections and /weights
 * [views][rows][channels], block means of projections and block sums of
 * weights, beam-hardening coefficient 0, BH_QUAD_COEF of
 * Standard_Data_Format/XT_Constants.h:8) and read_AngleTimeReconList
 * (XT_HDFIO.c:137-215; float /angles and /times, proj_num from
 * proj_start).  Unlike the reference, a dataset of the wrong rank is an
 * error (-1), not only a message. */
int32_t mbir_read_std_hdf5(const char* data_filename, float* projections, float* weights,
                           float* proj_angles, float* proj_times, int32_t datafile_row0,
                           int32_t proj_rows, int32_t proj_cols, int32_t proj_start,
                           int32_t proj_num, FILE* debug_file_ptr);

#ifdef __cplusplus
}
#endif
#endif
