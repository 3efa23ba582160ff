/*
 * pcmbir_engine.c -- host side of the PCMBIR drop-in: phcontomo_reconstruct()
 * (PCMBIR_3D/XT_Engine.c:82-196) over the phase-retrieval C-ABI of
 * include/pcmbir_hip.h.  Paths of reference files are relative to the
 * reference's Source_Code/.
 *
 * With the tomographic reconstruction compiled out of the reference
 * (ENABLE_TOMO_RECONS is never defined), a call runs one multi-resolution
 * stage (MAX_MULTRES_NUM = 1, PCMBIR_3D/XT_Constants.h:184) whose work is
 * either the Paganin retrieval of every view (recon_type 1,
 * do_PagPhaseRet_MBIRRecon, PCMBIR_3D/XT_ICD_update.c:536-600) or the ADMM
 * phase retrieval of every view started from it (recon_type 2,
 * PCMBIR_3D/XT_ICD_update.c:1020-1080), and writes the reference's files.
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "mbir_hip.h"
#include "pcmbir4d.h"
#include "pcmbir_hip.h"

/* PCMBIR_3D/XT_Constants.h */
#define PC_MIN_XY_RECON_RES 16        /* :183 */
#define PC_MAX_MULTRES_NUM 1          /* :184 */
#define PC_MIN_ROWS_PER_NODE 2        /* :185 */
#define PC_MIN_PROJECTION_ROWS 4      /* :186 */
#define PC_FORWPROJ_ADD_FRACTION 0.01 /* :87 */
#define PC_OBJECT_INIT_VAL 1e-15      /* :78-79 */
#define PC_PLANCKS_CONSTANT 4.135668e-19
#define PC_LIGHT_SPEED 299792458e6
#define PC_MRF_P 1.2                  /* :146 */
#define PC_PHANTOM_XY_SIZE 512        /* :203 */
#define PC_PHANTOM_Z_SIZE 128         /* :204 */
#define PC_EXPECTED_COUNT 2500.0      /* EXPECTED_COUNT_MEASUREMENT, :89 */
#define PC_NOISE_SEED 761521          /* srandom2 seed, PCMBIR_3D/XT_Engine.c:212 */
/* iteration limits, thresholds and ADMM weights (PCMBIR_3D/XT_Init.c:494-521) */
#define PC_PRET_MAXITER 100
#define PC_SD_MAXITER 100
#define PC_PRET_THRESH 0.0000001
#define PC_SD_THRESH 0.0005
#define PC_NU_TYPE2 0.5

static void say(FILE* fp, const char* tag, const char* fmt, ...) {
  va_list ap;
  if (!fp) return;
  fprintf(fp, "[%s] (pcmbir_engine) ", tag);
  va_start(ap, fmt);
  vfprintf(fp, fmt, ap);
  va_end(ap);
  fflush(fp);
}

static int write_bin(const char* name, const void* data, size_t elem, size_t n,
                     const char* mode, FILE* log) {
  char path[256];
  snprintf(path, sizeof path, "%s.bin", name);
  FILE* fp = fopen(path, mode);
  if (!fp) {
    say(log, "ERROR", "Cannot open file %s.\n", path);
    return -1;
  }
  size_t w = n ? fwrite(data, elem, n, fp) : 0;
  fclose(fp);
  if (w != n) say(log, "ERROR", "Short write to %s.\n", path);
  return w == n ? 0 : -1;
}

static int read_bin(const char* name, void* data, size_t elem, size_t n, FILE* log) {
  char path[256];
  snprintf(path, sizeof path, "%s.bin", name);
  FILE* fp = fopen(path, "rb");
  if (!fp) {
    say(log, "ERROR", "Error in reading file %s.\n", path);
    return -1;
  }
  size_t r = fread(data, elem, n, fp);
  fclose(fp);
  if (r != n) {
    say(log, "ERROR", "read_SharedBinFile_At: Read %zu number of elements from the file %s. "
        "However, required number of elements is %zu.\n", r, name, n);
    return -1;
  }
  return 0;
}

/* Zeros of n doubles, written in chunks (the reference's untouched arrays). */
static int write_zeros(const char* name, size_t n, FILE* log) {
  char path[256];
  snprintf(path, sizeof path, "%s.bin", name);
  FILE* fp = fopen(path, "wb");
  if (!fp) {
    say(log, "ERROR", "Cannot open file %s.\n", path);
    return -1;
  }
  static const double z[4096];
  size_t left = n;
  while (left) {
    size_t k = left < 4096 ? left : 4096;
    if (fwrite(z, sizeof(double), k, fp) != k) break;
    left -= k;
  }
  fclose(fp);
  return left ? -1 : 0;
}

/* Real or imaginary part of an interleaved complex float array, as doubles. */
static int write_part(const char* name, const float* c, size_t n, int imag, FILE* log) {
  double* d = (double*)malloc(sizeof(double) * (n ? n : 1));
  if (!d) {
    say(log, "ERROR", "Out of memory.\n");
    return -1;
  }
  for (size_t i = 0; i < n; ++i) d[i] = c[2 * i + imag];
  int rc = write_bin(name, d, sizeof(double), n, "wb", log);
  free(d);
  return rc;
}

static int is_pow2(int v) { return v >= 2 && v <= 1024 && (v & (v - 1)) == 0; }

/* The number of OpenMP threads the reference would size its z blocks from
 * (omp_get_max_threads: OMP_NUM_THREADS, else the online processors). */
static int reference_threads(void) {
  const char* e = getenv("OMP_NUM_THREADS");
  int t = e ? atoi(e) : 0;
  if (t <= 0) t = (int)sysconf(_SC_NPROCESSORS_ONLN);
  return t > 0 ? t : 1;
}

/* TomoAux entries [0..3] of one sinogram element, as initErrorSinogam sets
 * them for recon_type 2 (PCMBIR_3D/XT_ICD_update.c:741-776) from the initial
 * projection e and its average; slot 1 is overwritten by the retrieval. */
static void tomo_aux(const double* e, int N_r, int N_t, int j, int k, double avg,
                     double out[4]) {
  const double here = e[(size_t)j * N_t + k];
  out[1] = here;
  out[2] = PC_FORWPROJ_ADD_FRACTION * avg + (j - 1 >= 0 ? e[(size_t)(j - 1) * N_t + k] : here);
  out[3] = 2 * PC_FORWPROJ_ADD_FRACTION * avg +
           (j + 1 < N_r ? e[(size_t)(j + 1) * N_t + k] : 2 * here);
  out[0] = (out[1] + out[2]) / 2.0;
}

/* [N_p][N_r][N_t][4] doubles: slot 1 from z, the others from aux0 (initial
 * projection and average) or zero. */
static int write_tomo_aux(const char* name, const double* z, const double* aux0, double avg,
                          int N_p, int N_r, int N_t, FILE* log) {
  char path[256];
  snprintf(path, sizeof path, "%s.bin", name);
  FILE* fp = fopen(path, "wb");
  if (!fp) {
    say(log, "ERROR", "Cannot open file %s.\n", path);
    return -1;
  }
  const size_t per = (size_t)N_r * N_t;
  double* row = (double*)malloc(sizeof(double) * 4 * per);
  int rc = row ? 0 : -1;
  for (int i = 0; i < N_p && !rc; ++i) {
    for (int j = 0; j < N_r; ++j)
      for (int k = 0; k < N_t; ++k) {
        double* o = row + 4 * ((size_t)j * N_t + k);
        if (aux0) {
          tomo_aux(aux0 + (size_t)i * per, N_r, N_t, j, k, avg, o);
        } else {
          o[0] = o[2] = o[3] = 0.0;
        }
        o[1] = z[(size_t)i * per + (size_t)j * N_t + k];
      }
    if (fwrite(row, sizeof(double), 4 * per, fp) != 4 * per) rc = -1;
  }
  free(row);
  fclose(fp);
  if (rc) say(log, "ERROR", "Writing %s failed.\n", path);
  return rc;
}

static int write_scalar(const char* name, double v, const char* mode, FILE* log) {
  return write_bin(name, &v, sizeof v, 1, mode, log);
}

/* One multi-resolution stage (the only one, mult_xy = mult_z = 1). */
static int run_stage(const float* measurements, const float* brights, int32_t N_t, int32_t N_r,
                     int32_t N_p, int32_t recon_num, float vox_wid, float mag_sig_s,
                     float mag_c_s, float phase_sig_s, float phase_c_s, float obj2det_dist,
                     float light_energy, float pag_regparam, uint8_t recon_type, FILE* log) {
  const size_t per = (size_t)N_r * N_t, nall = per * (size_t)N_p;
  int rc = -1;
  pcmbir_hip_ctx* ctx = NULL;
  float *y = NULL, *D = NULL, *lam = NULL, *w = NULL, *om = NULL, *v = NULL, *win = NULL;
  double *zr = NULL, *zi = NULL, *er = NULL, *ei = NULL;
  double *c0 = NULL, *c1 = NULL, *pr = NULL, *du = NULL, *cs = NULL;
  int32_t* its = NULL;
  char name[128];

  if (!is_pow2(N_r) || !is_pow2(N_t)) {
    say(log, "ERROR", "proj_cols = %d and proj_rows = %d must be powers of two up to 1024 "
        "(pad the data; the GPU transforms are radix-2).\n", N_r, N_t);
    return -1;
  }
  /* XT_Init.c:256-259 (PCMBIR_3D): wavelength with the reference's constant */
  const double wavelength = PC_PLANCKS_CONSTANT * PC_LIGHT_SPEED / light_energy;
  const double wavenumber = 2 * M_PI / wavelength;
  const double delta = (double)vox_wid; /* Length_R / N_r = Length_T / N_t */
  y = (float*)malloc(sizeof(float) * 2 * nall);
  D = (float*)malloc(sizeof(float) * 2 * nall);
  lam = (float*)malloc(sizeof(float) * nall);
  win = (float*)malloc(sizeof(float) * per);
  zr = (double*)calloc(nall, sizeof(double));
  zi = (double*)calloc(nall, sizeof(double));
  if (!y || !D || !lam || !win || !zr || !zi) {
    say(log, "ERROR", "Out of memory.\n");
    goto done;
  }
  /* measurement amplitudes and D (PCMBIR_3D/XT_Init.c:362-389): the
   * sqrt-mean bright normalises both; the sinogram is [view][r][t] while the
   * inputs are [view][t][r] */
  {
    double bavg = 0.0;
    for (size_t q = 0; q < per; ++q) bavg += brights[q];
    bavg = sqrt(bavg / (double)per);
    for (int32_t i = 0; i < N_p; ++i)
      for (int32_t j = 0; j < N_r; ++j)
        for (int32_t k = 0; k < N_t; ++k) {
          const size_t o = (size_t)i * per + (size_t)j * N_t + k;
          const size_t in = (size_t)i * per + (size_t)k * N_r + j;
          y[2 * o] = (float)(sqrt((double)measurements[in]) / bavg);
          y[2 * o + 1] = 0.f;
          D[2 * o] = (float)(sqrt((double)brights[(size_t)k * N_r + j]) / bavg);
          D[2 * o + 1] = 0.f;
          lam[o] = 1.f;
        }
  }
  /* Fresnel window, sigma = 2 / sqrt(lambda z) (PCMBIR_3D/XT_Init.c:524-528,
   * create_FresnelTranWindow :210-231) */
  {
    const double sig = 2.0 / sqrt(wavelength * obj2det_dist);
    for (int32_t j = 0; j < N_r; ++j)
      for (int32_t k = 0; k < N_t; ++k) {
        const double u = (j >= N_r / 2 ? -(double)(N_r - j) : (double)j) / (N_r * delta);
        const double vv = (k >= N_t / 2 ? -(double)(N_t - k) : (double)k) / (N_t * delta);
        win[(size_t)j * N_t + k] = (float)exp(-(u * u + vv * vv) / (2 * sig * sig));
      }
  }
  if (pcmbir_hip_create(&ctx, N_p, N_r, N_t, delta, delta, wavelength, obj2det_dist, win)) {
    say(log, "ERROR", "GPU set-up failed: %s\n", pcmbir_hip_last_error());
    goto done;
  }
  /* residual files of the head loop and of each view, zero-initialised
   * (PCMBIR_3D/XT_ICD_update.c:970-980) */
  if (write_scalar("mag_recon_primal", 0.0, "wb", log) ||
      write_scalar("mag_recon_dual", 0.0, "wb", log) ||
      write_scalar("phase_recon_primal", 0.0, "wb", log) ||
      write_scalar("phase_recon_dual", 0.0, "wb", log))
    goto done;
  for (int32_t i = 0; i < N_p; ++i) {
    snprintf(name, sizeof name, "pret_primal_proj_%d", i);
    if (write_scalar(name, 0.0, "wb", log)) goto done;
    snprintf(name, sizeof name, "pret_dual_proj_%d", i);
    if (write_scalar(name, 0.0, "wb", log)) goto done;
  }

  w = (float*)calloc(2 * nall, sizeof(float));
  om = (float*)calloc(2 * nall, sizeof(float));
  v = (float*)calloc(2 * nall, sizeof(float));
  if (!w || !om || !v) {
    say(log, "ERROR", "Out of memory.\n");
    goto done;
  }
  double magavg = 0.0, phaseavg = 0.0;
  if (recon_type == 1) {
    if (pcmbir_hip_paganin(ctx, y, D, pag_regparam, wavenumber, zr, zi)) {
      say(log, "ERROR", "Paganin retrieval failed: %s\n", pcmbir_hip_last_error());
      goto done;
    }
    if (write_bin("pag_mag_ret", zr, sizeof(double), nall, "wb", log) ||
        write_bin("pag_phase_ret", zi, sizeof(double), nall, "wb", log))
      goto done;
  } else if (recon_type == 2) {
    er = (double*)malloc(sizeof(double) * nall);
    ei = (double*)malloc(sizeof(double) * nall);
    if (!er || !ei) {
      say(log, "ERROR", "Out of memory.\n");
      goto done;
    }
    /* initObject reads the Paganin estimate (PCMBIR_3D/XT_ObjectInit.c:425-431) */
    if (read_bin("pag_mag_ret", er, sizeof(double), nall, log) ||
        read_bin("pag_phase_ret", ei, sizeof(double), nall, log)) {
      say(log, "ERROR", "recon_type 2 starts from pag_mag_ret.bin / pag_phase_ret.bin "
          "(run recon_type 1 first).\n");
      goto done;
    }
    for (size_t q = 0; q < nall; ++q) {
      magavg += er[q];
      phaseavg += ei[q];
      zr[q] = er[q];
      zi[q] = ei[q];
    }
    magavg /= (double)nall;
    phaseavg /= (double)nall;
    say(log, "DEBUG", "Average of magnitude and phase components of froward projection in node 0 "
        "are %f and %f\n", magavg, phaseavg);
    /* w = exp(-z), v = 0, Omega = phase projection of w
     * (PCMBIR_3D/XT_ICD_update.c:778-789) */
    for (size_t q = 0; q < nall; ++q) {
      const double e = exp(-zr[q]);
      w[2 * q] = (float)(e * cos(-zi[q]));
      w[2 * q + 1] = (float)(e * sin(-zi[q]));
    }
    if (pcmbir_hip_set_phase_problem(ctx, y, D, lam, v, w, NULL, PC_NU_TYPE2) ||
        pcmbir_hip_set_nms_prior(ctx, phase_sig_s, phase_c_s, mag_sig_s, mag_c_s, pag_regparam,
                                 1, PC_MRF_P, PC_NU_TYPE2)) {
      say(log, "ERROR", "GPU set-up failed: %s\n", pcmbir_hip_last_error());
      goto done;
    }
    say(log, "INFO", "Doing phase retrieval ....\n");
    c0 = (double*)malloc(sizeof(double) * N_p);
    c1 = (double*)malloc(sizeof(double) * N_p);
    pr = (double*)malloc(sizeof(double) * N_p * PC_PRET_MAXITER);
    du = (double*)malloc(sizeof(double) * N_p * PC_PRET_MAXITER);
    cs = (double*)malloc(sizeof(double) * N_p * PC_PRET_MAXITER);
    its = (int32_t*)malloc(sizeof(int32_t) * N_p);
    if (!c0 || !c1 || !pr || !du || !cs || !its) {
      say(log, "ERROR", "Out of memory.\n");
      goto done;
    }
    if (pcmbir_hip_estimate(ctx, zr, zi, PC_PRET_MAXITER, PC_PRET_THRESH, PC_SD_MAXITER,
                            PC_SD_THRESH, its, c0, c1, pr, du, cs) ||
        pcmbir_hip_get_phase_state(ctx, w, om) || pcmbir_hip_get_dual(ctx, v)) {
      say(log, "ERROR", "Phase retrieval failed: %s\n", pcmbir_hip_last_error());
      goto done;
    }
    /* per view: pretcost history and residual histories
     * (PCMBIR_3D/XT_CmplxProjEst.c:120-121, 199-204) */
    for (int32_t i = 0; i < N_p; ++i) {
      const size_t o = (size_t)i * PC_PRET_MAXITER;
      snprintf(name, sizeof name, "pret_cost_proj_%d", i);
      if (write_bin(name, &c0[i], sizeof(double), 1, "wb", log) ||
          write_bin(name, cs + o, sizeof(double), (size_t)its[i], "ab", log))
        goto done;
      snprintf(name, sizeof name, "pret_primal_proj_%d", i);
      if (write_bin(name, pr + o, sizeof(double), (size_t)its[i], "ab", log)) goto done;
      snprintf(name, sizeof name, "pret_dual_proj_%d", i);
      if (write_bin(name, du + o, sizeof(double), (size_t)its[i], "ab", log)) goto done;
      say(log, "INFO", "**** PRet total iters is %d. Proj idx = %d. Primal and dual residual are "
          "%e and %e. Init cost is %e and new cost is %e. Total reduction in cost is %e ****\n",
          its[i], i, its[i] ? pr[o + its[i] - 1] : 0.0, its[i] ? du[o + its[i] - 1] : 0.0, c0[i],
          c1[i], c0[i] - c1[i]);
    }
  }

  /* tail files of ICD_BackProject (PCMBIR_3D/XT_ICD_update.c:1189-1204) */
  {
    int nzb = reference_threads() / recon_num;
    if (nzb < 2) nzb = 2;
    else if (nzb > N_t) nzb = N_t;
    nzb = (nzb / 2) * 2;
    if (write_zeros("update_map", (size_t)recon_num * nzb * (size_t)N_r * N_r, log)) goto done;
  }
  if (write_tomo_aux("mag_tomo_aux", zr, er, magavg, N_p, N_r, N_t, log) ||
      write_tomo_aux("phase_tomo_aux", zi, ei, phaseavg, N_p, N_r, N_t, log) ||
      write_part("mag_pret_aux", w, nall, 0, log) || write_part("phase_pret_aux", w, nall, 1, log) ||
      write_part("omega_real", om, nall, 0, log) || write_part("omega_imag", om, nall, 1, log) ||
      write_zeros("mag_tomo_dual", nall, log) || write_zeros("phase_tomo_dual", nall, log) ||
      write_part("mag_pret_dual", v, nall, 0, log) || write_part("phase_pret_dual", v, nall, 1, log))
    goto done;
  rc = 0;
done:
  pcmbir_hip_destroy(ctx);
  free(y);
  free(D);
  free(lam);
  free(w);
  free(om);
  free(v);
  free(win);
  free(zr);
  free(zi);
  free(er);
  free(ei);
  free(c0);
  free(c1);
  free(pr);
  free(du);
  free(cs);
  free(its);
  return rc;
}

int phcontomo_reconstruct(float** magobject, float** phaseobject, float* measurements,
                          float* brights, float* proj_angles, float* proj_times,
                          float* recon_times, int32_t proj_rows, int32_t proj_cols,
                          int32_t proj_num, int32_t recon_num, float vox_wid, float rot_center,
                          float mag_sig_s, float mag_sig_t, float mag_c_s, float mag_c_t,
                          float phase_sig_s, float phase_sig_t, float phase_c_s,
                          float phase_c_t, float convg_thresh, float obj2det_dist,
                          float light_energy, float pag_regparam, uint8_t restart,
                          uint8_t recon_type, FILE* debug_msg_ptr) {
  FILE* dbg = debug_msg_ptr;
  const time_t start = time(NULL);
  int32_t last_multres = 0;
  (void)proj_angles;
  (void)proj_times;
  (void)recon_times;
  (void)rot_center;
  (void)mag_sig_t;
  (void)mag_c_t;
  (void)phase_sig_t;
  (void)phase_c_t;
  (void)convg_thresh;
  say(dbg, "INFO", "Reconstructing the data ....\n");
  if (!magobject || !phaseobject || !measurements || !brights) {
    say(dbg, "ERROR", "Null input array.\n");
    return -1;
  }
  /* PCMBIR_3D/XT_Engine.c:100-103 with one process */
  if (proj_rows < PC_MIN_PROJECTION_ROWS) {
    say(dbg, "ERROR", "The minimum number of projection rows should be %d.\n",
        PC_MIN_PROJECTION_ROWS);
    return -1;
  }
  if (proj_cols % 2 != 0) {
    say(dbg, "ERROR", "The number of projection columns should be even.\n");
    return -1;
  }
  if (proj_rows % 2 != 0 || proj_rows < PC_MIN_ROWS_PER_NODE) {
    say(dbg, "ERROR", "The number of projection rows divided by the number of nodes should be an "
        "even number greater than or equal to %d.\n", PC_MIN_ROWS_PER_NODE);
    return -1;
  }
  if (proj_num <= 0 || recon_num <= 0) {
    say(dbg, "ERROR", "proj_num and recon_num must be positive.\n");
    return -1;
  }
  /* stage count (PCMBIR_3D/XT_Engine.c:110-115): at most MAX_MULTRES_NUM */
  int32_t multres_num = (int32_t)(log(((float)proj_cols) / PC_MIN_XY_RECON_RES) / log(2.0) + 1);
  if (multres_num < 2) multres_num = 2;
  if (multres_num > PC_MAX_MULTRES_NUM) multres_num = PC_MAX_MULTRES_NUM;
  if (restart == 1 && read_bin("status", &last_multres, sizeof last_multres, 1, dbg)) return -1;
  say(dbg, "INFO", "Number of multi-resolution stages is %d.\n", multres_num);
  for (int32_t i = last_multres; i < multres_num; ++i) {
    say(dbg, "INFO", "Running multi-resolution stage %d with x-y voxel scale = %d and z voxel "
        "scale = %d.\n", i, 1, 1);
    if (run_stage(measurements, brights, proj_rows, proj_cols, proj_num, recon_num, vox_wid,
                  mag_sig_s, mag_c_s, phase_sig_s, phase_c_s, obj2det_dist, light_energy,
                  pag_regparam, recon_type, dbg)) {
      say(dbg, "ERROR", "Reconstruction failed!\n");
      return -1;
    }
    say(dbg, "INFO", "Time elapsed is %f minutes.\n", difftime(time(NULL), start) / 60.0);
    /* object files (write_ObjectProjOff2TiffBinPerIter,
     * PCMBIR_3D/XT_IOMisc.c:236-265): the objects are never reconstructed;
     * recon_type 2 initialises them to 1e-15 (PCMBIR_3D/XT_ObjectInit.c:416-421) */
    {
      const size_t nvox = (size_t)proj_rows * proj_cols * proj_cols;
      double* obj = (double*)malloc(sizeof(double) * nvox);
      if (!obj) {
        say(dbg, "ERROR", "Out of memory.\n");
        return -1;
      }
      const double val = recon_type == 2 ? PC_OBJECT_INIT_VAL : 0.0;
      for (size_t q = 0; q < nvox; ++q) obj[q] = val;
      char name[64];
      int bad = 0;
      for (int32_t t = 0; t < recon_num && !bad; ++t) {
        snprintf(name, sizeof name, "mag_object_time_%d", t);
        bad = write_bin(name, obj, sizeof(double), nvox, "wb", dbg);
        snprintf(name, sizeof name, "phase_object_time_%d", t);
        bad = bad || write_bin(name, obj, sizeof(double), nvox, "wb", dbg);
      }
      free(obj);
      if (bad) return -1;
    }
    if (write_bin("status", &last_multres, sizeof last_multres, 1, "wb", dbg)) return -1;
    say(dbg, "INFO", "Completed multi-resolution stage %d.\n", i);
  }
  *magobject = NULL;
  *phaseobject = NULL;
  say(dbg, "INFO", "Exiting MBIR 4D\n");
  return 0;
}

/* Standard normal deviates for the measurement noise: the polar Box-Muller
 * method over splitmix64.  The reference draws them from its randlib
 * normal(); the sequence differs, the distribution is the same. */
typedef struct {
  uint64_t s;
  double spare;
  int has;
} pc_rng;

static double pc_uniform(pc_rng* r) {
  uint64_t z = (r->s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static double pc_normal(pc_rng* r) {
  if (r->has) {
    r->has = 0;
    return r->spare;
  }
  double u, v, q;
  do {
    u = 2.0 * pc_uniform(r) - 1.0;
    v = 2.0 * pc_uniform(r) - 1.0;
    q = u * u + v * v;
  } while (q >= 1.0 || q == 0.0);
  const double f = sqrt(-2.0 * log(q) / q);
  r->spare = v * f;
  r->has = 1;
  return u * f;
}

int phcontomo_forward_project_ex(float* measurements, float* brights, const float* mag,
                                 const float* phase, int32_t n_xy, int32_t n_z,
                                 const float* proj_angles, int32_t proj_rows, int32_t proj_cols,
                                 int32_t proj_num, float vox_wid, float rot_center,
                                 float obj2det_dist, float light_energy, uint64_t noise_seed,
                                 FILE* dbg) {
  if (!measurements || !brights || !mag || !phase || !proj_angles) {
    say(dbg, "ERROR", "phcontomo_forward_project: null argument.\n");
    return -1;
  }
  /* the reference's argument checks (PCMBIR_3D/XT_Engine.c:214-217, one node) */
  if (proj_rows < PC_MIN_PROJECTION_ROWS || proj_rows % 2 || proj_cols % 2 || proj_num < 1) {
    say(dbg, "ERROR", "proj_rows must be even and >= %d, proj_cols even.\n",
        PC_MIN_PROJECTION_ROWS);
    return -1;
  }
  if (n_xy % proj_cols || n_z % proj_rows || !is_pow2(n_xy) || !is_pow2(n_z)) {
    say(dbg, "ERROR", "phantom %d x %d: sizes must be powers of two up to 1024 (the GPU "
        "transforms are radix-2) and multiples of proj_cols / proj_rows.\n", n_xy, n_z);
    return -1;
  }
  const int32_t N_r = n_xy, N_t = n_z, rs = n_xy / proj_cols, ts = n_z / proj_rows;
  const size_t per = (size_t)N_r * N_t, nall = per * proj_num;
  /* XT_Init.c:616-619: wavelength, wavenumber; :623-628, 679: the sampling */
  const double wavelength = PC_PLANCKS_CONSTANT * PC_LIGHT_SPEED / light_energy;
  const double wavenumber = 2 * M_PI / wavelength;
  const double d_r = (double)vox_wid * proj_cols / N_r, d_t = (double)vox_wid * proj_rows / N_t;
  int rc = -1;
  float *pr = NULL, *pi = NULL, *field = NULL, *win = NULL;
  pcmbir_hip_ctx* ctx = NULL;
  pr = (float*)malloc(sizeof(float) * nall);
  pi = (float*)malloc(sizeof(float) * nall);
  field = (float*)malloc(sizeof(float) * 2 * nall);
  win = (float*)malloc(sizeof(float) * per);
  if (!pr || !pi || !field || !win) {
    say(dbg, "ERROR", "Out of memory.\n");
    goto done;
  }
  /* projections of both phantoms (XT_genSinogram.c:146-186) */
  if (mbir_hip_phantom_project(mag, phase, n_xy, n_z, proj_num, proj_angles, proj_rows,
                               proj_cols, vox_wid, rot_center, wavenumber, pr, pi)) {
    say(dbg, "ERROR", "Phantom projection failed: %s\n", mbir_hip_last_error());
    goto done;
  }
  if (write_bin("SimMagProjs", pr, sizeof(float), nall, "wb", dbg) ||
      write_bin("SimPhaseProjs", pi, sizeof(float), nall, "wb", dbg) ||
      write_bin("mag_phantom", mag, sizeof(float), (size_t)n_z * n_xy * n_xy, "wb", dbg) ||
      write_bin("phase_phantom", phase, sizeof(float), (size_t)n_z * n_xy * n_xy, "wb", dbg))
    goto done;
  /* exit field sqrt(I0) exp(-p_mag - i p_phase) in the transform's [r][t]
   * layout (:214-228) */
  {
    const double amp = sqrt(PC_EXPECTED_COUNT);
    for (int32_t i = 0; i < proj_num; ++i)
      for (int32_t t = 0; t < N_t; ++t)
        for (int32_t r = 0; r < N_r; ++r) {
          const size_t src = (size_t)i * per + (size_t)t * N_r + r;
          const size_t dst = (size_t)i * per + (size_t)r * N_t + t;
          const double e = exp(-(double)pr[src]);
          field[2 * dst] = (float)(e * cos(-(double)pi[src]) * amp);
          field[2 * dst + 1] = (float)(e * sin(-(double)pi[src]) * amp);
        }
  }
  /* Fresnel propagation with the simulator's window, sigma = 1 /
   * sqrt(lambda z / 2) (XT_Init.c:682-686, create_FresnelTranWindow
   * :210-231) */
  {
    const double sig = 1.0 / sqrt(0.5 * wavelength * obj2det_dist);
    for (int32_t j = 0; j < N_r; ++j)
      for (int32_t k = 0; k < N_t; ++k) {
        const double u = (j >= N_r / 2 ? -(double)(N_r - j) : (double)j) / (N_r * d_r);
        const double v = (k >= N_t / 2 ? -(double)(N_t - k) : (double)k) / (N_t * d_t);
        win[(size_t)j * N_t + k] = (float)exp(-(u * u + v * v) / (2 * sig * sig));
      }
  }
  if (pcmbir_hip_create(&ctx, proj_num, N_r, N_t, d_r, d_t, wavelength, obj2det_dist, win) ||
      pcmbir_hip_fresnel(ctx, 0, field, field)) {
    say(dbg, "ERROR", "Fresnel transform failed: %s\n", pcmbir_hip_last_error());
    goto done;
  }
  /* intensity averaged over each detector pixel's ts x rs samples, plus
   * the noise |m + sqrt(|m|) n| (:241-264); the brights are the expected
   * count */
  {
    pc_rng rng = {noise_seed, 0.0, 0};
    for (int32_t i = 0; i < proj_num; ++i)
      for (int32_t row = 0; row < proj_rows; ++row)
        for (int32_t col = 0; col < proj_cols; ++col) {
          double m = 0.0;
          for (int32_t a = 0; a < ts; ++a)
            for (int32_t b = 0; b < rs; ++b) {
              const size_t q = (size_t)i * per + (size_t)(col * rs + b) * N_t + row * ts + a;
              const double re = field[2 * q], im = field[2 * q + 1];
              m += re * re + im * im;
            }
          m /= (double)(ts * rs);
          if (noise_seed) m = fabs(m + sqrt(fabs(m)) * pc_normal(&rng));
          measurements[((size_t)i * proj_rows + row) * proj_cols + col] = (float)m;
        }
    for (size_t q = 0; q < (size_t)proj_rows * proj_cols; ++q) brights[q] = (float)PC_EXPECTED_COUNT;
  }
  rc = 0;
done:
  if (ctx) pcmbir_hip_destroy(ctx);
  free(pr);
  free(pi);
  free(field);
  free(win);
  return rc;
}

int phcontomo_forward_project(float** measurements, float** brights, float* proj_angles,
                              int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                              float vox_wid, float rot_center, float obj2det_dist,
                              float light_energy, float pag_regparam, FILE* debug_msg_ptr) {
  (void)pag_regparam; /* stored by the reference, unused by its simulator */
  if (!measurements || !brights) return -1;
  /* the phantoms the reference reads from MAG_PHANTOM_FILEPATH /
   * PHASE_PHANTOM_FILEPATH (PCMBIR_3D/XT_Constants.h:91-92), here from the
   * files named by PHCONTOMO_MAG_PHANTOM / PHCONTOMO_PHASE_PHANTOM */
  const char* mpath = getenv("PHCONTOMO_MAG_PHANTOM");
  const char* ppath = getenv("PHCONTOMO_PHASE_PHANTOM");
  const size_t nv = (size_t)PC_PHANTOM_Z_SIZE * PC_PHANTOM_XY_SIZE * PC_PHANTOM_XY_SIZE;
  float* mag = (float*)malloc(sizeof(float) * nv);
  float* phase = (float*)malloc(sizeof(float) * nv);
  int rc = -1;
  FILE* fp = NULL;
  if (!mag || !phase || !mpath || !ppath) {
    say(debug_msg_ptr, "ERROR", "Set PHCONTOMO_MAG_PHANTOM and PHCONTOMO_PHASE_PHANTOM to the "
        "%d x %d x %d float phantom files.\n", PC_PHANTOM_Z_SIZE, PC_PHANTOM_XY_SIZE,
        PC_PHANTOM_XY_SIZE);
    goto done;
  }
  fp = fopen(mpath, "rb");
  if (!fp || fread(mag, sizeof(float), nv, fp) != nv) {
    say(debug_msg_ptr, "ERROR", "Error in reading file %s\n", mpath);
    goto done;
  }
  fclose(fp);
  fp = fopen(ppath, "rb");
  if (!fp || fread(phase, sizeof(float), nv, fp) != nv) {
    say(debug_msg_ptr, "ERROR", "Error in reading file %s\n", ppath);
    goto done;
  }
  rc = phcontomo_forward_project_ex(*measurements, *brights, mag, phase, PC_PHANTOM_XY_SIZE,
                                    PC_PHANTOM_Z_SIZE, proj_angles, proj_rows, proj_cols,
                                    proj_num, vox_wid, rot_center, obj2det_dist, light_energy,
                                    PC_NOISE_SEED, debug_msg_ptr);
done:
  if (fp) fclose(fp);
  free(mag);
  free(phase);
  return rc;
}
