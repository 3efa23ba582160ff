/*
 * mbir_hdf5.c -- host-side HDF5 input stages of the real-data drivers
 * (include/mbir_hdf5.h): the hyperslabs the drivers' readers select, read
 * with libhdf5, then normalised and down-sampled on the GPU by libmbir4d.
 * One process: the reference splits the detector rows over MPI ranks
 * (reconstruct/real_data/APS_Data_Format/XT_HDFIO.c:71-72,97-103); here the
 * rank count is one.
 */
#include "mbir_hdf5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdlib.h>

#include "mbir_hip.h"

static void msg(FILE* log, const char* fmt, ...) {
  if (!log) return;
  va_list ap;
  va_start(ap, fmt);
  vfprintf(log, fmt, ap);
  va_end(ap);
  fflush(log);
}

/* An open dataset of a known rank and its extent. */
typedef struct {
  hid_t file, dset, space;
  int rank;
  hsize_t dims[3];
} h5in;

static void h5in_close(h5in* d) {
  if (d->space >= 0) H5Sclose(d->space);
  if (d->dset >= 0) H5Dclose(d->dset);
  if (d->file >= 0) H5Fclose(d->file);
  d->space = d->dset = d->file = -1;
}

static int h5in_open(h5in* d, const char* path, const char* name, int rank, FILE* log) {
  d->file = d->dset = d->space = -1;
  d->rank = 0;
  d->file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (d->file < 0) {
    msg(log, "ERROR: cannot open HDF5 file %s\n", path);
    return -1;
  }
  d->dset = H5Dopen2(d->file, name, H5P_DEFAULT);
  if (d->dset < 0) {
    msg(log, "ERROR: no dataset %s in %s\n", name, path);
    h5in_close(d);
    return -1;
  }
  d->space = H5Dget_space(d->dset);
  d->rank = H5Sget_simple_extent_ndims(d->space);
  if (d->rank != rank) {
    msg(log, "ERROR: The rank of the dataset %s is not %d.\n", name, rank);
    h5in_close(d);
    return -1;
  }
  H5Sget_simple_extent_dims(d->space, d->dims, NULL);
  return 0;
}

/* Reads the block at off of size cnt (rank d->rank) as type into out. */
static int h5in_read(h5in* d, const hsize_t* off, const hsize_t* cnt, hid_t type, void* out,
                     FILE* log) {
  if (H5Sselect_hyperslab(d->space, H5S_SELECT_SET, off, NULL, cnt, NULL) < 0) {
    msg(log, "ERROR: hyperslab selection failed.\n");
    return -1;
  }
  hid_t mem = H5Screate_simple(d->rank, cnt, NULL);
  const herr_t st = H5Dread(d->dset, type, mem, d->space, H5P_DEFAULT, out);
  H5Sclose(mem);
  if (st < 0) {
    msg(log, "ERROR: reading the HDF5 dataset failed.\n");
    return -1;
  }
  return 0;
}

int32_t mbir_read_aps_hdf5(const char* data_filename, const char* whites_filename,
                           const char* darks_filename, float* projections, float* weights,
                           int32_t datafile_row0, int32_t proj_rows, int32_t proj_cols,
                           int32_t proj_start, int32_t proj_num, FILE* log) {
  h5in P, Wh, Dk;
  uint16_t *pbuf = NULL, *wbuf = NULL, *dbuf = NULL;
  int32_t rc = -1;
  P.file = Wh.file = Dk.file = -1;
  P.dset = Wh.dset = Dk.dset = -1;
  P.space = Wh.space = Dk.space = -1;
  if (!data_filename || !whites_filename || !darks_filename || !projections || !weights ||
      proj_rows < 1 || proj_cols < 1 || proj_num < 1 || proj_start < 0 || datafile_row0 < 0) {
    msg(log, "ERROR: bad argument to mbir_read_aps_hdf5.\n");
    return -1;
  }
  if (h5in_open(&P, data_filename, "/exchange/data", 3, log) ||
      h5in_open(&Wh, whites_filename, "/exchange/data_white", 3, log) ||
      h5in_open(&Dk, darks_filename, "/exchange/data_dark", 3, log))
    goto out;
  msg(log, "Size of white (/exchange/data_white) dataset is %dx%dx%d\n", (int)Wh.dims[0],
      (int)Wh.dims[1], (int)Wh.dims[2]);
  msg(log, "Size of dark (/exchange/data_dark) dataset is %dx%dx%d\n", (int)Dk.dims[0],
      (int)Dk.dims[1], (int)Dk.dims[2]);
  msg(log, "Size of count (/exchange/data) dataset is %dx%dx%d\n", (int)P.dims[0],
      (int)P.dims[1], (int)P.dims[2]);
  /* the channel extents must agree and hold proj_cols (XT_HDFIO.c:53-63) */
  if (Wh.dims[2] != P.dims[2] || Dk.dims[2] != P.dims[2] || P.dims[2] < (hsize_t)proj_cols) {
    msg(log, "ERROR: channel counts: white %d, dark %d, data %d, proj_cols = %d\n",
        (int)Wh.dims[2], (int)Dk.dims[2], (int)P.dims[2], proj_cols);
    goto out;
  }
  {
    /* the down-sampling ratio from the channels, the same along the rows;
     * the remainder channels split off both sides (XT_HDFIO.c:65-70, 87-95) */
    const hsize_t extra = P.dims[2] % (hsize_t)proj_cols;
    const hsize_t width = P.dims[2] - extra;
    const int32_t ratio = (int32_t)(width / (hsize_t)proj_cols);
    const hsize_t slices = (hsize_t)proj_rows * (hsize_t)ratio;
    const hsize_t need = (hsize_t)datafile_row0 + slices;
    if (Wh.dims[1] != P.dims[1] || Dk.dims[1] != P.dims[1] || P.dims[1] < need) {
      msg(log, "ERROR: rows: white %d, dark %d, data %d, need at least %d\n", (int)Wh.dims[1],
          (int)Dk.dims[1], (int)P.dims[1], (int)need);
      goto out;
    }
    if (Wh.dims[0] < 3 || Dk.dims[0] < 3) {
      msg(log, "ERROR: white and dark sets need at least 3 frames (the first and last are "
               "not used).\n");
      goto out;
    }
    if ((hsize_t)proj_start + (hsize_t)proj_num > P.dims[0]) {
      msg(log, "ERROR: views %d .. %d exceed the %d frames of /exchange/data\n", proj_start,
          proj_start + proj_num - 1, (int)P.dims[0]);
      goto out;
    }
    const hsize_t nw = Wh.dims[0] - 2, nd = Dk.dims[0] - 2;
    const size_t frame = (size_t)slices * (size_t)width;
    pbuf = (uint16_t*)malloc(sizeof(uint16_t) * frame * (size_t)proj_num);
    wbuf = (uint16_t*)malloc(sizeof(uint16_t) * frame * (size_t)nw);
    dbuf = (uint16_t*)malloc(sizeof(uint16_t) * frame * (size_t)nd);
    if (!pbuf || !wbuf || !dbuf) {
      msg(log, "ERROR: out of memory.\n");
      goto out;
    }
    /* white and dark frames 1 .. N-2, the views from proj_start
     * (XT_HDFIO.c:78-96) */
    const hsize_t woff[3] = {1, (hsize_t)datafile_row0, extra / 2};
    const hsize_t wcnt[3] = {nw, slices, width};
    const hsize_t dcnt[3] = {nd, slices, width};
    const hsize_t poff[3] = {(hsize_t)proj_start, (hsize_t)datafile_row0, extra / 2};
    const hsize_t pcnt[3] = {(hsize_t)proj_num, slices, width};
    msg(log, "Reading HDF5 dataset ...\n");
    if (h5in_read(&Wh, woff, wcnt, H5T_NATIVE_UINT16, wbuf, log) ||
        h5in_read(&Dk, woff, dcnt, H5T_NATIVE_UINT16, dbuf, log) ||
        h5in_read(&P, poff, pcnt, H5T_NATIVE_UINT16, pbuf, log))
      goto out;
    msg(log, "Data sub-sampling factor along x-axis = %d, data sub-sampling factor along "
             "z-axis = %d\n", ratio, ratio);
    if (mbir_hip_aps_normalize(pbuf, wbuf, dbuf, proj_num, (int32_t)nw, (int32_t)nd, proj_rows,
                               proj_cols, ratio, 0.5, projections, weights)) {
      msg(log, "ERROR: GPU normalisation failed: %s\n", mbir_hip_last_error());
      goto out;
    }
    msg(log, "Generated projections and weight data with beamhardening coefficient of %f\n",
        0.5);
    rc = 0;
  }
out:
  free(pbuf);
  free(wbuf);
  free(dbuf);
  h5in_close(&P);
  h5in_close(&Wh);
  h5in_close(&Dk);
  return rc;
}

int32_t mbir_read_std_hdf5(const char* data_filename, float* projections, float* weights,
                           float* proj_angles, float* proj_times, int32_t datafile_row0,
                           int32_t proj_rows, int32_t proj_cols, int32_t proj_start,
                           int32_t proj_num, FILE* log) {
  h5in Pj, Wt, An, Tm;
  float *pbuf = NULL, *wbuf = NULL;
  int32_t rc = -1;
  Pj.file = Wt.file = An.file = Tm.file = -1;
  Pj.dset = Wt.dset = An.dset = Tm.dset = -1;
  Pj.space = Wt.space = An.space = Tm.space = -1;
  if (!data_filename || !projections || !weights || !proj_angles || !proj_times ||
      proj_rows < 1 || proj_cols < 1 || proj_num < 1 || proj_start < 0 || datafile_row0 < 0) {
    msg(log, "ERROR: bad argument to mbir_read_std_hdf5.\n");
    return -1;
  }
  if (h5in_open(&Wt, data_filename, "/weights", 3, log) ||
      h5in_open(&Pj, data_filename, "/projections", 3, log))
    goto out;
  msg(log, "Size of weights (/weights) dataset is %dx%dx%d\n", (int)Wt.dims[0],
      (int)Wt.dims[1], (int)Wt.dims[2]);
  msg(log, "Size of projections (/projections) dataset is %dx%dx%d\n", (int)Pj.dims[0],
      (int)Pj.dims[1], (int)Pj.dims[2]);
  if (Wt.dims[0] != Pj.dims[0] || Wt.dims[1] != Pj.dims[1] || Wt.dims[2] != Pj.dims[2]) {
    msg(log, "ERROR: Dimensions of weights dataset and projection datasets don't match\n");
    goto out;
  }
  if (Pj.dims[2] < (hsize_t)proj_cols) {
    msg(log, "ERROR: %d channels hold fewer than proj_cols = %d\n", (int)Pj.dims[2], proj_cols);
    goto out;
  }
  {
    /* ratio and centred channels as the APS reader; the block must lie in
     * the datasets (Standard_Data_Format/XT_HDFIO.c:50-74) */
    const hsize_t extra = Pj.dims[2] % (hsize_t)proj_cols;
    const hsize_t width = Pj.dims[2] - extra;
    const int32_t ratio = (int32_t)(width / (hsize_t)proj_cols);
    const hsize_t slices = (hsize_t)proj_rows * (hsize_t)ratio;
    const hsize_t off[3] = {(hsize_t)proj_start, (hsize_t)datafile_row0, extra / 2};
    const hsize_t cnt[3] = {(hsize_t)proj_num, slices, width};
    if (off[0] + cnt[0] > Pj.dims[0] || off[1] + cnt[1] > Pj.dims[1] ||
        off[2] + cnt[2] > Pj.dims[2]) {
      msg(log, "ERROR: Dataset size is inconsistent with inputs\n");
      goto out;
    }
    const size_t n = (size_t)cnt[0] * cnt[1] * cnt[2];
    pbuf = (float*)malloc(sizeof(float) * n);
    wbuf = (float*)malloc(sizeof(float) * n);
    if (!pbuf || !wbuf) {
      msg(log, "ERROR: out of memory.\n");
      goto out;
    }
    msg(log, "Data sub-sampling factor along x-axis = %d, data sub-sampling factor along "
             "z-axis = %d\n", ratio, ratio);
    if (h5in_read(&Wt, off, cnt, H5T_NATIVE_FLOAT, wbuf, log) ||
        h5in_read(&Pj, off, cnt, H5T_NATIVE_FLOAT, pbuf, log))
      goto out;
    if (mbir_hip_std_downsample(pbuf, wbuf, proj_num, proj_rows, proj_cols, ratio, 0.0,
                                projections, weights)) {
      msg(log, "ERROR: GPU down-sampling failed: %s\n", mbir_hip_last_error());
      goto out;
    }
    msg(log, "Generated projections and weight data with beamhardening coefficient of %f\n",
        0.0);
  }
  /* the view angles and times (Standard_Data_Format/XT_HDFIO.c:137-215) */
  if (h5in_open(&An, data_filename, "/angles", 1, log) ||
      h5in_open(&Tm, data_filename, "/times", 1, log))
    goto out;
  if (An.dims[0] != Tm.dims[0]) {
    msg(log, "ERROR: Size of angles and times list does not match\n");
    goto out;
  }
  if ((hsize_t)proj_start + (hsize_t)proj_num > An.dims[0]) {
    msg(log, "ERROR: Dataset size is inconsistent with inputs\n");
    goto out;
  }
  {
    const hsize_t off1[1] = {(hsize_t)proj_start}, cnt1[1] = {(hsize_t)proj_num};
    if (h5in_read(&An, off1, cnt1, H5T_NATIVE_FLOAT, proj_angles, log) ||
        h5in_read(&Tm, off1, cnt1, H5T_NATIVE_FLOAT, proj_times, log))
      goto out;
  }
  rc = 0;
out:
  free(pbuf);
  free(wbuf);
  h5in_close(&Pj);
  h5in_close(&Wt);
  h5in_close(&An);
  h5in_close(&Tm);
  return rc;
}
