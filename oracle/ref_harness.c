/*
 * ref_harness.c -- thin ctypes-callable wrapper around the reference
 * MBIR_4D library, compiled together with the reference's own C sources
 * (read in place under the reference tree; see oracle/Makefile) into
 * oracle/_ref/libmbir4d_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, tests/golden/make_golden.py and
 * bench.py's cpu_baseline leg to run the reference's functions on the same
 * inputs as the HIP path.  Nothing here re-implements reference arithmetic;
 * every number comes from a reference function:
 *   initStructures             XT_Init.c:211
 *   DetectorResponseProfile    XT_Profile.c:129
 *   calcAMatrixColumnforAngle  XT_AMatrix.c:50
 *   initErrorSinogam           XT_ICD_update.c:651
 *   CE_FunctionalSubstitution  XT_ICD_update.c:133
 *   updateVoxels               XT_VoxUpdate.c:286
 *   computeCost                XT_ICD_update.c:185
 *   updateVoxelsTimeSlices     XT_ICD_update.c:932
 *   reconstruct                XT_Engine.c:81
 *
 * Sinogram-shaped arrays cross this interface in the reference's internal
 * [view][r][t] order; objects as [time][z][y][x] without halo slices.
 */
#include <mpi.h>
#include <omp.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "XT_Structures.h"
#include "XT_AMatrix.h"
#include "XT_ICD_update.h"
#include "XT_Init.h"
#include "XT_NHICD.h"
#include "XT_Profile.h"
#include "XT_VoxUpdate.h"
#include "allocate.h"
#include "randlib.h"

/* Reference functions that are non-static but not declared in its headers. */
int32_t initErrorSinogam(Sinogram*, ScannedObject*, TomoInputs*, Real_arr_t**,
                         Real_arr_t***);
Real_t computeCost(Sinogram*, ScannedObject*, TomoInputs*, Real_arr_t***);
int updateVoxelsTimeSlices(Sinogram*, ScannedObject*, TomoInputs*,
                           Real_arr_t**, Real_arr_t***, int32_t,
                           Real_arr_t****, uint8_t***);
void update_Sinogram_Offset(Sinogram*, TomoInputs*, Real_arr_t***);
void update_variance_parameter(Sinogram*, TomoInputs*, Real_arr_t***);
void gen_offset_constraint_windows(Sinogram*, TomoInputs*);
void ComputeVSC(ScannedObject*, Real_arr_t**, Real_arr_t**, Real_arr_t**);
int reconstruct(float**, float*, float*, float*, float*, float*, int32_t,
                int32_t, int32_t, int32_t, float, float, float, float, float,
                float, float, int32_t, int32_t, float, float, uint8_t, FILE*);

typedef struct {
  Sinogram S;
  ScannedObject O;
  TomoInputs T;
  Real_arr_t** H_r;
  Real_arr_t* H_t;
  Real_arr_t*** esino;
  Real_arr_t**** magmap;
  uint8_t*** mask;
  float *proj, *wgt, *angles, *ptimes, *rtimes;
  FILE* log;
  int windows;
} ref_ctx;

static void ref_mpi_up(void) {
  int up = 0;
  MPI_Initialized(&up);
  if (!up) MPI_Init(NULL, NULL);
}

static float* ref_dup(const float* a, size_t n) {
  float* b = (float*)malloc(sizeof(float) * (n ? n : 1));
  memcpy(b, a, sizeof(float) * n);
  return b;
}

/* Builds the reference's structures for one multi-resolution stage with the
 * given voxel scales (mult_idx 0: object initialised to zero, no files
 * read).  The TIFF dumps initStructures always writes go to scratch_dir. */
void* ref_create(const char* scratch_dir, int32_t mult_xy, int32_t mult_z,
                 const float* projections, const float* weights,
                 const float* angles, const float* ptimes,
                 const float* rtimes, int32_t rows, int32_t cols, int32_t num,
                 int32_t recon_num, float vox_wid, float rot_center,
                 float sig_s, float sig_t, float c_s, float c_t, float convg,
                 int32_t remove_rings, int32_t quad_convex, float huber_delta,
                 float huber_T, int32_t threads) {
  ref_mpi_up();
  if (threads > 0) omp_set_num_threads(threads);
  ref_ctx* c = (ref_ctx*)calloc(1, sizeof(ref_ctx));
  size_t ns = (size_t)num * rows * cols;
  c->proj = ref_dup(projections, ns);
  c->wgt = ref_dup(weights, ns);
  c->angles = ref_dup(angles, num);
  c->ptimes = ref_dup(ptimes, num);
  c->rtimes = ref_dup(rtimes, recon_num + 1);
  c->log = fopen("/dev/null", "w");
  c->T.debug_file_ptr = c->log;
  int32_t mxy[1] = {mult_xy}, mz[1] = {mult_z};
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) cwd[0] = 0;
  if (scratch_dir && chdir(scratch_dir) != 0) return NULL;
  int32_t flag = initStructures(&c->S, &c->O, &c->T, 0, mxy, mz, c->proj,
                                c->wgt, c->angles, c->ptimes, c->rtimes, rows,
                                cols, num, recon_num, vox_wid, rot_center,
                                sig_s, sig_t, c_s, c_t, convg, remove_rings,
                                quad_convex, huber_delta, huber_T);
  if (scratch_dir && cwd[0] && chdir(cwd) != 0) return NULL;
  if (flag) return NULL;
  c->T.Write2Tiff = 0;
  srandom2(761521); /* XT_Engine.c:99 */
  c->H_r = (Real_arr_t**)multialloc(sizeof(Real_arr_t), 2, c->S.N_p,
                                    DETECTOR_RESPONSE_BINS + 1);
  c->H_t = (Real_arr_t*)get_spc(DETECTOR_RESPONSE_BINS + 1, sizeof(Real_arr_t));
  DetectorResponseProfile(c->H_r, c->H_t, &c->S, &c->O, &c->T);
  c->esino = (Real_arr_t***)multialloc(sizeof(Real_arr_t), 3, c->S.N_p,
                                       c->S.N_r, c->S.N_t);
  memset(&c->esino[0][0][0], 0, sizeof(Real_arr_t) * c->S.N_p * c->S.N_r * c->S.N_t);
  return c;
}

/* Geometry as the reference computed it: ints[8], dbls[24]. */
void ref_geometry(void* h, int32_t* ints, double* dbls) {
  ref_ctx* c = (ref_ctx*)h;
  int i, j, k, n = 0;
  ints[0] = c->S.N_r; ints[1] = c->S.N_t; ints[2] = c->S.N_p;
  ints[3] = c->O.N_x; ints[4] = c->O.N_y; ints[5] = c->O.N_z;
  ints[6] = c->O.N_time; ints[7] = c->T.num_z_blocks;
  dbls[n++] = c->S.delta_r; dbls[n++] = c->S.delta_t; dbls[n++] = c->S.R0;
  dbls[n++] = c->S.RMax; dbls[n++] = c->S.T0; dbls[n++] = c->S.TMax;
  dbls[n++] = c->S.OffsetR; dbls[n++] = c->S.OffsetT; dbls[n++] = c->O.x0;
  dbls[n++] = c->O.y0; dbls[n++] = c->O.z0; dbls[n++] = c->O.delta_xy;
  dbls[n++] = c->O.delta_z; dbls[n++] = c->T.radius_obj;
  dbls[n++] = c->O.delta_recon; dbls[n++] = c->T.Sigma_S_Q;
  dbls[n++] = c->T.Sigma_S_Q_P; dbls[n++] = c->T.Sigma_T_Q;
  dbls[n++] = c->T.Sigma_T_Q_P; dbls[n++] = c->T.Time_Filter[0];
  dbls[n++] = c->T.alpha; dbls[n++] = c->T.ErrorSinoThresh;
  dbls[n++] = c->T.ErrorSinoDelta; dbls[n++] = c->T.var_est;
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j)
      for (k = 0; k < 3; ++k) dbls[n++] = c->T.Spatial_Filter[i][j][k];
}

/* View lists per frame: ptr[N_time+1], views[N_p]. */
void ref_view_map(void* h, int32_t* ptr, int32_t* views) {
  ref_ctx* c = (ref_ctx*)h;
  int32_t n = 0, i, k;
  for (i = 0; i < c->O.N_time; ++i) {
    ptr[i] = n;
    for (k = 0; k < c->O.ProjNum[i]; ++k) views[n++] = c->O.ProjIdxPtr[i][k];
  }
  ptr[c->O.N_time] = n;
}

void ref_detector_response(void* h, float* H) {
  ref_ctx* c = (ref_ctx*)h;
  memcpy(H, &c->H_r[0][0], sizeof(float) * c->S.N_p * (DETECTOR_RESPONSE_BINS + 1));
}

int32_t ref_amatrix_column(void* h, int32_t row, int32_t col, int32_t view,
                           int32_t* idx, double* val) {
  ref_ctx* c = (ref_ctx*)h;
  AMatrixCol a;
  int32_t m;
  a.values = (Real_t*)get_spc(256, sizeof(Real_t));
  a.index = (int32_t*)get_spc(256, sizeof(int32_t));
  calcAMatrixColumnforAngle(&c->S, &c->O, c->H_r, &a, row, col, view);
  for (m = 0; m < a.count; ++m) {
    idx[m] = a.index[m];
    val[m] = a.values[m];
  }
  m = a.count;
  free(a.values);
  free(a.index);
  return m;
}

void ref_set_object(void* h, const float* x) {
  ref_ctx* c = (ref_ctx*)h;
  int32_t t, z, Nz = c->O.N_z, plane = c->O.N_y * c->O.N_x;
  for (t = 0; t < c->O.N_time; ++t) {
    memset(&c->O.Object[t][0][0][0], 0, sizeof(float) * plane);
    memset(&c->O.Object[t][Nz + 1][0][0], 0, sizeof(float) * plane);
    for (z = 0; z < Nz; ++z)
      memcpy(&c->O.Object[t][z + 1][0][0], x + ((size_t)t * Nz + z) * plane,
             sizeof(float) * plane);
  }
}

void ref_get_object(void* h, float* x) {
  ref_ctx* c = (ref_ctx*)h;
  int32_t t, z, Nz = c->O.N_z, plane = c->O.N_y * c->O.N_x;
  for (t = 0; t < c->O.N_time; ++t)
    for (z = 0; z < Nz; ++z)
      memcpy(x + ((size_t)t * Nz + z) * plane, &c->O.Object[t][z + 1][0][0],
             sizeof(float) * plane);
}

/* e = y - A x - d from the current object (callable once per context: the
 * reference frees the projections at the end). */
int32_t ref_init_error(void* h, float* e) {
  ref_ctx* c = (ref_ctx*)h;
  if (initErrorSinogam(&c->S, &c->O, &c->T, c->H_r, c->esino)) return -1;
  memcpy(e, &c->esino[0][0][0], sizeof(float) * c->S.N_p * c->S.N_r * c->S.N_t);
  return 0;
}

void ref_get_error(void* h, float* e) {
  ref_ctx* c = (ref_ctx*)h;
  memcpy(e, &c->esino[0][0][0], sizeof(float) * c->S.N_p * c->S.N_r * c->S.N_t);
}

void ref_get_weight(void* h, float* w) {
  ref_ctx* c = (ref_ctx*)h;
  memcpy(w, &c->T.Weight[0][0][0], sizeof(float) * c->S.N_p * c->S.N_r * c->S.N_t);
}

double ref_substitution(void* h, double V, double th1, double th2,
                        const double* nb, const uint8_t* nbf, const double* tn,
                        const uint8_t* tf) {
  ref_ctx* c = (ref_ctx*)h;
  Real_t sn[3][3][3], t[2];
  bool sf[3][3][3], tb[2];
  int i, j, k;
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j)
      for (k = 0; k < 3; ++k) {
        sn[i][j][k] = nb[(i * 3 + j) * 3 + k];
        sf[i][j][k] = nbf[(i * 3 + j) * 3 + k] != 0;
      }
  t[0] = tn[0]; t[1] = tn[1];
  tb[0] = tf[0] != 0; tb[1] = tf[1] != 0;
  return CE_FunctionalSubstitution(V, th1, th2, &c->O, &c->T, sn, t, sf, tb);
}

/* Updates one voxel through the reference's voxel-line driver (one line,
 * one slice).  Returns the number of zero-skipped voxels (0 or 1). */
long ref_update_voxel(void* h, int32_t t, int32_t z, int32_t row, int32_t col,
                      int32_t iter) {
  ref_ctx* c = (ref_ctx*)h;
  int32_t xs = col, ys = row;
  long zc = 0;
  Real_arr_t** mag = (Real_arr_t**)multialloc(sizeof(Real_arr_t), 2, c->O.N_y, c->O.N_x);
  uint8_t** m = (uint8_t**)multialloc(sizeof(uint8_t), 2, c->O.N_y, c->O.N_x);
  memset(&m[0][0], 1, c->O.N_y * c->O.N_x);
  updateVoxels(t, t, z, z, 0, 0, &xs, &ys, &c->S, &c->O, &c->T, c->esino,
               c->H_r, iter, &zc, mag, m);
  multifree(mag, 2);
  multifree(m, 2);
  return zc;
}

double ref_cost(void* h) {
  ref_ctx* c = (ref_ctx*)h;
  return computeCost(&c->S, &c->O, &c->T, c->esino);
}

void ref_cost_parts(void* h, double* out) {
  ref_ctx* c = (ref_ctx*)h;
  out[0] = c->T.Forward_Cost;
  out[1] = c->T.Prior_Cost;
}

void ref_srandom2(uint32_t seed) { srandom2(seed); }
double ref_random2(void) { return random2(); }

/* Runs n reference ICD iterations (HICD + NHICD schedule) numbered from
 * first_iter, as ICD_BackProject's loop does (XT_ICD_update.c:1127-1172):
 * costs[i] after each iteration, conv[i] = updateVoxelsTimeSlices' flag,
 * secs[i] = wall seconds of the update.  init_error must have run. */
int32_t ref_icd_run(void* h, int32_t n, int32_t first_iter, double* costs,
                    int32_t* conv, double* secs) {
  ref_ctx* c = (ref_ctx*)h;
  int32_t i, j, k, it;
  if (!c->mask) {
    c->mask = (uint8_t***)multialloc(sizeof(uint8_t), 3, c->O.N_time, c->O.N_y, c->O.N_x);
    for (i = 0; i < c->O.N_time; i++)
      for (j = 0; j < c->O.N_y; j++)
        for (k = 0; k < c->O.N_x; k++) {
          Real_t x = c->O.x0 + ((Real_t)k + 0.5) * c->O.delta_xy;
          Real_t y = c->O.y0 + ((Real_t)j + 0.5) * c->O.delta_xy;
          c->mask[i][j][k] = (x * x + y * y < c->T.radius_obj * c->T.radius_obj) ? 1 : 0;
        }
    c->magmap = (Real_arr_t****)multialloc(sizeof(Real_arr_t), 4, c->O.N_time,
                                           c->T.num_z_blocks, c->O.N_y, c->O.N_x);
    memset(&c->magmap[0][0][0][0], 0,
           sizeof(Real_arr_t) * c->O.N_time * c->T.num_z_blocks * c->O.N_y * c->O.N_x);
  }
  for (it = 0; it < n; ++it) {
    double t0 = omp_get_wtime();
    conv[it] = updateVoxelsTimeSlices(&c->S, &c->O, &c->T, c->H_r, c->esino,
                                      first_iter + it, c->magmap, c->mask);
    secs[it] = omp_get_wtime() - t0;
    costs[it] = computeCost(&c->S, &c->O, &c->T, c->esino);
  }
  return 0;
}

/* The reference's full multi-resolution reconstruct(), run in scratch_dir
 * (it writes its stage files and TIFF dumps there).  object receives
 * [recon_num][rows+2][cols][cols] floats (halo slices included). */
int32_t ref_reconstruct(const char* scratch_dir, float* object,
                        const float* projections, const float* weights,
                        const float* angles, const float* ptimes,
                        const float* rtimes, int32_t rows, int32_t cols,
                        int32_t num, int32_t recon_num, float vox_wid,
                        float rot_center, float sig_s, float sig_t, float c_s,
                        float c_t, float convg, int32_t remove_rings,
                        int32_t quad_convex, float huber_delta, float huber_T,
                        int32_t threads) {
  float* obj = NULL;
  char cwd[4096];
  int32_t flag;
  size_t ns = (size_t)num * rows * cols;
  float *p = ref_dup(projections, ns), *w = ref_dup(weights, ns);
  float *a = ref_dup(angles, num), *pt = ref_dup(ptimes, num);
  float* rt = ref_dup(rtimes, recon_num + 1);
  FILE* log = fopen("/dev/null", "w");
  ref_mpi_up();
  if (threads > 0) omp_set_num_threads(threads);
  if (!getcwd(cwd, sizeof cwd)) return -1;
  if (chdir(scratch_dir) != 0) return -1;
  flag = reconstruct(&obj, p, w, a, pt, rt, rows, cols, num, recon_num,
                     vox_wid, rot_center, sig_s, sig_t, c_s, c_t, convg,
                     remove_rings, quad_convex, huber_delta, huber_T, 0, log);
  if (chdir(cwd) != 0) flag = -1;
  if (flag == 0 && obj)
    memcpy(object, obj, sizeof(float) * (size_t)recon_num * (rows + 2) * cols * cols);
  free(obj);
  free(p); free(w); free(a); free(pt); free(rt);
  fclose(log);
  return flag;
}

/* Sinogram-sized state in and out of the reference's arrays ([p][r][t]). */
static void ref_put_sino(ref_ctx* c, const float* w, const float* e, const uint8_t* sel) {
  const size_t n = (size_t)c->S.N_p * c->S.N_r * c->S.N_t;
  for (size_t i = 0; i < n; ++i) {
    if (w) (&c->T.Weight[0][0][0])[i] = w[i];
    (&c->esino[0][0][0])[i] = e[i];
    (&c->S.ProjSelect[0][0][0])[i] = sel[i] != 0;
  }
}

static void ref_get_sino(ref_ctx* c, float* w, float* e, uint8_t* sel) {
  const size_t n = (size_t)c->S.N_p * c->S.N_r * c->S.N_t;
  for (size_t i = 0; i < n; ++i) {
    if (w) w[i] = (&c->T.Weight[0][0][0])[i];
    if (e) e[i] = (&c->esino[0][0][0])[i];
    sel[i] = (&c->S.ProjSelect[0][0][0])[i] ? 1 : 0;
  }
}

/* One ring-offset update (update_Sinogram_Offset) of the given mode on the
 * given weights, error sinogram, offset and selection.  Mode 3's windows
 * come from gen_offset_constraint_windows, which the reference calls only
 * when it simulates data (XT_Init.c:528), not on its reconstruction path. */
void ref_offset_update(void* h, int32_t mode, const float* w, float* e, float* d, uint8_t* sel) {
  ref_ctx* c = (ref_ctx*)h;
  ref_put_sino(c, w, e, sel);
  memcpy(&c->S.ProjOffset[0][0], d, sizeof(float) * c->S.N_r * c->S.N_t);
  c->T.OffsetConstraintType = mode;
  if (mode == 3 && !c->windows) {
    gen_offset_constraint_windows(&c->S, &c->T);
    c->windows = 1;
  }
  update_Sinogram_Offset(&c->S, &c->T, c->esino);
  ref_get_sino(c, NULL, e, sel);
  memcpy(d, &c->S.ProjOffset[0][0], sizeof(float) * c->S.N_r * c->S.N_t);
}

/* The variance re-estimate (update_variance_parameter) from var on the
 * given weights (already divided by var), error sinogram and selection;
 * returns the new estimate, w and sel receive the updated arrays. */
double ref_variance_update(void* h, double var, float* w, const float* e, uint8_t* sel) {
  ref_ctx* c = (ref_ctx*)h;
  ref_put_sino(c, w, e, sel);
  c->T.var_est = var;
  update_variance_parameter(&c->S, &c->T, c->esino);
  ref_get_sino(c, w, NULL, sel);
  return c->T.var_est;
}

/* ComputeVSC (XT_NHICD.c:126-160) of one magnitude map [N_y][N_x] into
 * out [N_y][N_x]. */
void ref_vsc(void* h, const float* mag, float* out) {
  ref_ctx* c = (ref_ctx*)h;
  const size_t n = (size_t)c->O.N_y * c->O.N_x;
  Real_arr_t** m = (Real_arr_t**)multialloc(sizeof(Real_arr_t), 2, c->O.N_y, c->O.N_x);
  Real_arr_t** f = (Real_arr_t**)multialloc(sizeof(Real_arr_t), 2, c->O.N_y, c->O.N_x);
  Real_arr_t** f2 = (Real_arr_t**)multialloc(sizeof(Real_arr_t), 2, c->O.N_y, c->O.N_x);
  memcpy(&m[0][0], mag, sizeof(float) * n);
  ComputeVSC(&c->O, m, f, f2);
  memcpy(out, &f[0][0], sizeof(float) * n);
  multifree(m, 2);
  multifree(f, 2);
  multifree(f2, 2);
}

/* VSC_based_Voxel_Line_Select (XT_NHICD.c:164-216) on the magnitude maps
 * mag [N_time][N_y][N_x] (the same map in every z-block) with
 * UpdateSelectNum = count lines per frame and block.  sel [N_time][N_y][N_x]
 * receives 1 for every line z-block 0's lists select; returns the number
 * selected in frame 0. */
int32_t ref_vsc_select(void* h, const float* mag, int32_t count, uint8_t* sel) {
  ref_ctx* c = (ref_ctx*)h;
  const int32_t nt = c->O.N_time, nb = c->T.num_z_blocks, ny = c->O.N_y, nx = c->O.N_x;
  Real_arr_t**** m = (Real_arr_t****)multialloc(sizeof(Real_arr_t), 4, nt, nb, ny, nx);
  for (int32_t i = 0; i < nt; i++)
    for (int32_t b = 0; b < nb; b++) {
      memcpy(&m[i][b][0][0], mag + (size_t)i * ny * nx, sizeof(float) * ny * nx);
      c->T.UpdateSelectNum[i][b] = count;
    }
  VSC_based_Voxel_Line_Select(&c->O, &c->T, m);
  memset(sel, 0, (size_t)nt * ny * nx);
  for (int32_t i = 0; i < nt; i++)
    for (int32_t k = 0; k < c->T.NHICDSelectNum[i][0]; k++)
      sel[((size_t)i * ny + c->T.y_NHICD_select[i][0][k]) * nx + c->T.x_NHICD_select[i][0][k]] = 1;
  multifree(m, 4);
  return c->T.NHICDSelectNum[0][0];
}

void ref_destroy(void* h) {
  ref_ctx* c = (ref_ctx*)h;
  if (!c) return;
  if (c->log) fclose(c->log);
  free(c);
}
