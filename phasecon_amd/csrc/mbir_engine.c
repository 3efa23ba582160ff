/*
 * mbir_engine.c -- C host of the MBIR engine: the drop-in reconstruct()
 * (include/mbir4d.h) and its stage set-up, driving the HIP kernels through
 * the thin C-ABI of include/mbir_hip.h.
 *
 * Reference behaviour restated here (paths relative to the reference's
 * Source_Code/ directory):
 *   multi-resolution schedule and stage loop   MBIR_4D/XT_Engine.c:81-194
 *   stage geometry, prior weights, view map    MBIR_4D/XT_Init.c:56-432
 *   detector response table                    MBIR_4D/XT_Profile.c:47-218
 *   object initialisation / up-sampling        MBIR_4D/XT_ICD_update.c:358-647
 *   ICD outer loop, convergence test, files    MBIR_4D/XT_ICD_update.c:1095-1230
 *   output files                               MBIR_4D/XT_IOMisc.c:60-108,237-269
 * The hot path (A build, A x, the ICD sweeps, the FP64 cost) runs on the
 * GPU; this file only prepares inputs and sequences the stages.
 */
#include "mbir4d.h"
#include "mbir_engine.h"
#include "mbir_hip.h"

#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#define RESP_BINS 64     /* DETECTOR_RESPONSE_BINS, XT_Constants.h:79 */
#define PROFILE_RES 1536 /* PROFILE_RESOLUTION, XT_Constants.h:77 */
#define BEAM_RES 512     /* BEAM_RESOLUTION, XT_Constants.h:78 */
#define PI_D 3.14159265358979323846
#define PI_2_D 1.57079632679489661923132169163975144
#define PI_4_D 0.785398163397448309615660845819875721
#define NHICD_SUBSETS 10 /* NHICD_Iterations, XT_Init.c:406 */

static void say(FILE* fp, const char* tag, const char* fmt, ...) {
  va_list ap;
  if (!fp) return;
  fprintf(fp, "[%s] (mbir_engine) ", tag);
  va_start(ap, fmt);
  vfprintf(fp, fmt, ap);
  va_end(ap);
  fflush(fp);
}

/* ------------------------------------------------------------ schedule */
int32_t mbir_multires_schedule(int32_t proj_rows, int32_t proj_cols,
                               int32_t* mult_xy, int32_t* mult_z) {
  int32_t fine_xy[MBIR_MAX_STAGES], fine_z[MBIR_MAX_STAGES];
  int32_t n = (int32_t)(log(((float)proj_cols) / 64) / log(2.0) + 1);
  if (n > MBIR_MAX_STAGES) n = MBIR_MAX_STAGES;
  if (n < 2) n = 2;
  fine_xy[0] = 1;
  fine_z[0] = 1;
  for (int32_t i = 1; i < n; ++i) {
    if (proj_cols % (fine_xy[i - 1] * 2) != 0) {
      n = i;
      break;
    }
    fine_xy[i] = fine_xy[i - 1] * 2;
    const int32_t z2 = fine_z[i - 1] * 2;
    fine_z[i] = (proj_rows % z2 == 0 && proj_rows / z2 >= 2) ? z2 : fine_z[i - 1];
  }
  for (int32_t i = 0; i < n; ++i) {
    mult_xy[i] = fine_xy[n - 1 - i];
    mult_z[i] = fine_z[n - 1 - i];
  }
  return n;
}

int32_t mbir_interlaced_views(float* views, float* times, int32_t K, int32_t N_theta,
                              int32_t proj_start, int32_t N_p, float min_acq_time,
                              float rot_speed) {
  if (!views || !times || K < 1 || N_theta < K || proj_start < 0 || N_p < 1) return -1;
  const int32_t L = N_theta / K;
  const int32_t bits = (int32_t)floor(log((double)K) / log(2.0) + 0.5);
  const float dtheta = (float)(M_PI / N_theta);
  const float min_step = rot_speed * min_acq_time;
  const int32_t total = proj_start + N_p;
  float* ang = (float*)malloc(sizeof(float) * (size_t)total);
  float* tim = (float*)malloc(sizeof(float) * (size_t)total);
  if (!ang || !tim) {
    free(ang);
    free(tim);
    return -1;
  }
  int32_t kept = 0;
  for (int32_t i = 0; kept < total; ++i) {
    const int32_t sub = (i / L) % K;
    int32_t rev = 0;
    for (int32_t b = 0; b < bits; ++b)
      if (sub & (1 << b)) rev |= 1 << (bits - 1 - b);
    const float index = (float)(i * K + rev);
    ang[kept] = index * dtheta;
    tim[kept] = index / (float)K;
    if (kept == 0 || ang[kept] - ang[kept - 1] > min_step) ++kept;
  }
  memcpy(views, ang + proj_start, sizeof(float) * (size_t)N_p);
  memcpy(times, tim + proj_start, sizeof(float) * (size_t)N_p);
  free(ang);
  free(tim);
  return total;
}

int32_t mbir_view_map(int32_t N_p, const float* proj_times, int32_t N_time,
                      const float* recon_times, int32_t* frame_ptr,
                      int32_t* frame_views) {
  int32_t n = 0, resume = 0;
  for (int32_t v = 0; v + 1 < N_p; ++v)
    if (proj_times[v + 1] < proj_times[v]) return -1;
  for (int32_t f = 0; f < N_time; ++f)
    if (recon_times[f + 1] < recon_times[f]) return -1;
  for (int32_t f = 0; f < N_time; ++f) {
    frame_ptr[f] = n;
    for (int32_t v = resume; v < N_p; ++v) {
      const double t = proj_times[v];
      if (t >= recon_times[f + 1]) {
        resume = v;
        break;
      }
      if (t >= recon_times[f]) frame_views[n++] = v;
    }
  }
  frame_ptr[N_time] = n;
  return n;
}

/* 1/|offset| over the 3x3x3 (+time) neighbourhood, XT_Init.c:138-141. */
static double nb_distance(int dt, int dz, int dy, int dx) {
  return sqrt(pow((double)dt - 1.0, 2.0) + pow((double)dz - 1.0, 2.0) +
              pow((double)dy - 1.0, 2.0) + pow((double)dx - 1.0, 2.0));
}

int mbir_stage_geometry(mbir_hip_geom* g, int32_t proj_rows,
                        int32_t proj_cols, int32_t proj_num, int32_t recon_num,
                        float vox_wid, float rot_center, float sig_s,
                        float sig_t, float c_s, float c_t, int32_t mult_xy,
                        int32_t mult_z, int32_t quad_convex, float huber_delta,
                        float huber_T, const int32_t* frame_ptr, double p_s,
                        double p_t) {
  const double vw = vox_wid, rc = rot_center;
  const double len_r = vw * proj_cols, len_t = vw * proj_rows;
  memset(g, 0, sizeof(*g));
  if (mult_xy <= 0 || mult_z <= 0 || recon_num <= 0) return -1;
  g->N_p = proj_num;
  g->N_r = proj_cols;
  g->N_t = proj_rows;
  g->N_time = recon_num;
  g->mult_xy = mult_xy;
  g->mult_z = mult_z;
  g->delta_r = len_r / g->N_r;
  g->delta_t = len_t / g->N_t;
  g->R0 = -rc * g->delta_r;
  g->RMax = (g->N_r - rc) * g->delta_r;
  g->N_x = (int32_t)(g->N_r / (double)mult_xy);
  g->N_y = g->N_x;
  g->N_z = (int32_t)(g->N_t / (double)mult_z);
  g->delta_xy = mult_xy * g->delta_r;
  g->x0 = g->R0;
  g->y0 = -len_r / 2.0;
  g->OffsetR = (g->delta_xy / sqrt(2.0) + g->delta_r / 2.0) / RESP_BINS;
  /* prior: sigma scaled by mult_xy (space) and the mean views per frame
   * (time); weights 1/distance normalised over the 26 + 2 neighbours and
   * scaled by mult_xy^3 * mean views per frame (XT_Init.c:126,149-173,
   * 418-421) */
  double views = 0;
  for (int32_t f = 0; f < recon_num; ++f) views += frame_ptr[f + 1] - frame_ptr[f];
  const double per_frame = views / recon_num;
  g->q = 2.0;
  g->p_s = p_s;
  g->p_t = p_t;
  g->c_s = c_s;
  g->c_t = c_t;
  g->sig_s_q = pow((double)sig_s * (double)mult_xy, g->q);
  g->sig_s_qp = pow((double)sig_s * (double)mult_xy, g->q - p_s);
  g->sig_t_q = pow((double)sig_t * per_frame, g->q);
  g->sig_t_qp = pow((double)sig_t * per_frame, g->q - p_t);
  const double scale = (double)mult_xy * mult_xy * mult_xy * per_frame;
  const double wt = 1.0 / nb_distance(0, 1, 1, 1);
  double norm = 2.0 * wt;
  for (int k = 0; k < 27; ++k) {
    const int a = k / 9, b = (k / 3) % 3, c = k % 3;
    g->filt[k] = 0.0;
    if (k != 13) {
      g->filt[k] = 1.0 / nb_distance(1, a, b, c);
      norm = norm + g->filt[k];
    }
  }
  for (int k = 0; k < 27; ++k) g->filt[k] = scale * g->filt[k] / norm;
  g->filt_t = scale * wt / norm;
  g->alpha = 1.5; /* OVER_RELAXATION_FACTOR, XT_Constants.h:87 */
  g->huber_T = quad_convex ? 100000 : huber_T;     /* XT_Constants.h:97 */
  g->huber_delta = quad_convex ? 1 : huber_delta;  /* XT_Constants.h:98 */
  g->var_est = 1.0;
  if (g->N_t % mult_z != 0 || g->N_r % mult_xy != 0) return -1;
  return 0;
}

void mbir_sincos(int32_t n, const float* angles, double* cosv, double* sinv) {
  for (int32_t i = 0; i < n; ++i) {
    cosv[i] = cos(angles[i]);
    sinv[i] = sin(angles[i]);
  }
}

void mbir_detector_response(const mbir_hip_geom* g, const float* angles,
                            float* H) {
  /* Voxel line-integral profile (a trapezoid over 2*delta_xy sampled at
   * 1536 points) convolved with a unit-area Hamming beam over one detector
   * pitch (512 samples), tabulated at 64 offsets of OffsetR per view. */
  double beam[BEAM_RES], bsum = 0;
  float* trap = (float*)malloc(sizeof(float) * PROFILE_RES);
  for (int i = 0; i < BEAM_RES; ++i) {
    beam[i] = 0.54 - 0.46 * cos((2 * PI_D / BEAM_RES) * i);
    bsum = bsum + beam[i];
  }
  for (int i = 0; i < BEAM_RES; ++i) beam[i] /= bsum;
  const double pitch = g->delta_r;
  const double first = -pitch / 2, step = pitch / BEAM_RES;
  const double to_bin = (PROFILE_RES) / (2 * g->delta_xy);
  for (int32_t v = 0; v < g->N_p; ++v) {
    double ang = angles[v];
    while (ang > PI_2_D) ang -= PI_2_D;
    while (ang < 0) ang += PI_2_D;
    const double top = (ang <= PI_4_D) ? g->delta_xy / cos(ang)
                                       : g->delta_xy / cos(PI_2_D - ang);
    const double h = cos(PI_4_D);
    const double outer = h * cos((PI_4_D - ang));
    const double inner = h * fabs((cos((PI_4_D + ang))));
    const double e0 = 1 - outer, e1 = 1 - inner, e2 = 1 + inner, e3 = 1 + outer;
    for (int j = 0; j < PROFILE_RES; ++j) {
      const double t = 2.0 * j / PROFILE_RES;
      double val;
      if (t <= e0 || t >= e3)
        val = 0;
      else if (t > e2)
        val = top * (e3 - t) / (e3 - e2);
      else if (t >= e1)
        val = top;
      else
        val = top * (t - e0) / (e1 - e0);
      trap[j] = (float)val;
    }
    float* row = H + (size_t)v * (RESP_BINS + 1);
    for (int i = 0; i < RESP_BINS; ++i) {
      const double lo = i * g->OffsetR - g->delta_xy;
      double acc = 0;
      for (int p = 0; p < BEAM_RES; ++p) {
        const double r = first + p * step;
        const int32_t b = (int32_t)floor((r - lo) * to_bin);
        if (b >= 0 && b <= PROFILE_RES - 1) acc += (trap[b] * beam[p]);
      }
      row[i] = (float)acc;
    }
    row[RESP_BINS] = 0.0f;
  }
  free(trap);
}

void mbir_update_mask(const mbir_hip_geom* g, double radius, uint8_t* mask) {
  for (int32_t j = 0; j < g->N_y; ++j)
    for (int32_t k = 0; k < g->N_x; ++k) {
      const double x = g->x0 + ((double)k + 0.5) * g->delta_xy;
      const double y = g->y0 + ((double)j + 0.5) * g->delta_xy;
      mask[(size_t)j * g->N_x + k] = (x * x + y * y < radius * radius) ? 1 : 0;
    }
}

void mbir_sino_to_internal(int32_t N_p, int32_t N_t, int32_t N_r,
                           const float* in, float* out) {
  for (int32_t p = 0; p < N_p; ++p)
    for (int32_t t = 0; t < N_t; ++t)
      for (int32_t r = 0; r < N_r; ++r)
        out[((size_t)p * N_r + r) * N_t + t] = in[((size_t)p * N_t + t) * N_r + r];
}

/* cost-rise guard of the drop-in's equit loop: relative FP32 noise
 * tolerance and recoveries before the reference's error */
#define GUARD_TOL 1e-6
#define GUARD_RECOVERIES 2

void mbir_default_options(mbir_options* o) {
  o->p_s = 1.2; /* MRF_P, XT_Constants.h:62 */
  o->p_t = 1.2;
  o->tile = 8;
  o->alpha = -1.0;
  o->max_equits = 1000; /* MAX_NUM_ITERATIONS, XT_Constants.h:86 */
  o->single_stage = 0;
  o->write_files = 1;
  o->stop_cost = 0.0;
  o->seed = 761521; /* XT_Engine.c:99 */
  o->refresh_every = 10;
}

/* ------------------------------------------------------- cost guard */
void mbir_guard_init(mbir_guard* gd, double cost0, double tol, int32_t max_recoveries) {
  gd->best = cost0;
  gd->tol = tol;
  gd->recoveries = 0;
  gd->max_recoveries = max_recoveries;
}

int32_t mbir_guard_check(mbir_guard* gd, double cost, int32_t exact) {
  /* NaN and infinity are rises whatever the tolerance */
  const int finite = cost == cost && cost < HUGE_VAL && cost > -HUGE_VAL;
  if (finite && cost <= gd->best + gd->tol * fabs(gd->best)) {
    if (cost < gd->best) {
      gd->best = cost;
      return MBIR_GUARD_SAVE;
    }
    return MBIR_GUARD_OK;
  }
  if (!exact && finite) return MBIR_GUARD_CONFIRM;
  if (gd->recoveries < gd->max_recoveries) {
    gd->recoveries++;
    return MBIR_GUARD_RECOVER;
  }
  return MBIR_GUARD_FAIL;
}

double mbir_guard_alpha(double alpha) { return alpha > 1.0 ? 1.0 + 0.5 * (alpha - 1.0) : 0.5 * alpha; }

/* ------------------------------------------------------------ files */
static int write_bin(const char* name, const void* data, size_t elem,
                     size_t n, const char* mode, FILE* log) {
  char path[256];
  snprintf(path, sizeof path, "%s.bin", name);
  FILE* fp = fopen(path, mode);
  if (!fp) {
    say(log, "ERROR", "Cannot open file %s.\n", path);
    return -1;
  }
  size_t w = fwrite(data, elem, n, fp);
  fclose(fp);
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
    say(log, "ERROR", "Number of elements read does not match required, required = %zu, read = %zu\n", n, r);
    return -1;
  }
  return 0;
}

/* ------------------------------------------------------- up-sampling */
/* One x-y plane [ny][nx] -> [2ny][2nx] (XT_ICD_update.c:358-390): ends
 * copied, interior samples at 1/4 and 3/4 between neighbours; the
 * intermediate row buffer is float as in the reference. */
static void upsample_plane(const float* in, int ny, int nx, float* out) {
  float* buf = (float*)malloc(sizeof(float) * (size_t)ny * 2 * nx);
  const int W2 = 2 * nx;
  for (int j = 0; j < ny; ++j) {
    const float* I = in + (size_t)j * nx;
    float* b = buf + (size_t)j * W2;
    b[0] = I[0];
    b[1] = (3.0 * I[0] + I[1]) / 4.0;
    b[W2 - 1] = I[nx - 1];
    b[W2 - 2] = (I[nx - 2] + 3.0 * I[nx - 1]) / 4.0;
    for (int k = 1; k < nx - 1; ++k) {
      b[2 * k] = (I[k - 1] + 3.0 * I[k]) / 4.0;
      b[2 * k + 1] = (3.0 * I[k] + I[k + 1]) / 4.0;
    }
  }
  for (int k = 0; k < W2; ++k) {
    out[k] = buf[k];
    out[(size_t)W2 + k] = (3.0 * buf[k] + buf[(size_t)W2 + k]) / 4.0;
    out[(size_t)(2 * ny - 1) * W2 + k] = buf[(size_t)(ny - 1) * W2 + k];
    out[(size_t)(2 * ny - 2) * W2 + k] =
        (buf[(size_t)(ny - 2) * W2 + k] + 3.0 * buf[(size_t)(ny - 1) * W2 + k]) / 4.0;
  }
  for (int j = 1; j < ny - 1; ++j)
    for (int k = 0; k < W2; ++k) {
      out[(size_t)(2 * j) * W2 + k] = (buf[(size_t)(j - 1) * W2 + k] + 3.0 * buf[(size_t)j * W2 + k]) / 4.0;
      out[(size_t)(2 * j + 1) * W2 + k] = (3 * buf[(size_t)j * W2 + k] + buf[(size_t)(j + 1) * W2 + k]) / 4.0;
    }
  free(buf);
}

/* [nz][ny][nx] -> [2nz][2ny][2nx] with FP64 intermediates
 * (XT_ICD_update.c:448-496). */
static void upsample_volume(const float* in, int nz, int ny, int nx, float* out) {
  const int W2 = 2 * nx, H2 = 2 * ny;
  const size_t plane = (size_t)H2 * W2;
  double* b2 = (double*)malloc(sizeof(double) * (size_t)nz * ny * W2);
  double* b3 = (double*)malloc(sizeof(double) * (size_t)nz * plane);
  for (int s = 0; s < nz; ++s) {
    for (int j = 0; j < ny; ++j) {
      const float* I = in + ((size_t)s * ny + j) * nx;
      double* b = b2 + ((size_t)s * ny + j) * W2;
      b[0] = I[0];
      b[1] = (3.0 * I[0] + I[1]) / 4.0;
      b[W2 - 1] = I[nx - 1];
      b[W2 - 2] = (I[nx - 2] + 3.0 * I[nx - 1]) / 4.0;
      for (int k = 1; k < nx - 1; ++k) {
        b[2 * k] = (I[k - 1] + 3.0 * I[k]) / 4.0;
        b[2 * k + 1] = (3.0 * I[k] + I[k + 1]) / 4.0;
      }
    }
    const double* B = b2 + (size_t)s * ny * W2;
    double* C = b3 + (size_t)s * plane;
    for (int k = 0; k < W2; ++k) {
      C[k] = B[k];
      C[(size_t)W2 + k] = (3.0 * B[k] + B[(size_t)W2 + k]) / 4.0;
      C[(size_t)(H2 - 1) * W2 + k] = B[(size_t)(ny - 1) * W2 + k];
      C[(size_t)(H2 - 2) * W2 + k] = (B[(size_t)(ny - 2) * W2 + k] + 3.0 * B[(size_t)(ny - 1) * W2 + k]) / 4.0;
    }
    for (int j = 1; j < ny - 1; ++j)
      for (int k = 0; k < W2; ++k) {
        C[(size_t)(2 * j) * W2 + k] = (B[(size_t)(j - 1) * W2 + k] + 3.0 * B[(size_t)j * W2 + k]) / 4.0;
        C[(size_t)(2 * j + 1) * W2 + k] = (3 * B[(size_t)j * W2 + k] + B[(size_t)(j + 1) * W2 + k]) / 4.0;
      }
  }
  for (size_t i = 0; i < plane; ++i) {
    out[i] = (float)b3[i];
    out[plane + i] = (float)((3.0 * b3[i] + b3[plane + i]) / 4.0);
    out[(size_t)(2 * nz - 1) * plane + i] = (float)b3[(size_t)(nz - 1) * plane + i];
    out[(size_t)(2 * nz - 2) * plane + i] =
        (float)((3.0 * b3[(size_t)(nz - 1) * plane + i] + b3[(size_t)(nz - 2) * plane + i]) / 4.0);
  }
  for (int s = 1; s < nz - 1; ++s)
    for (size_t i = 0; i < plane; ++i) {
      out[(size_t)(2 * s) * plane + i] =
          (float)((b3[(size_t)(s - 1) * plane + i] + 3.0 * b3[(size_t)s * plane + i]) / 4.0);
      out[(size_t)(2 * s + 1) * plane + i] =
          (float)((3.0 * b3[(size_t)s * plane + i] + b3[(size_t)(s + 1) * plane + i]) / 4.0);
    }
  free(b2);
  free(b3);
}

/* --------------------------------------------------------- one stage */
typedef struct {
  mbir_hip_geom g;
  int32_t* frame_ptr;
  int32_t* frame_views;
  float* H;
  double *cosv, *sinv;
  float *y, *w;
  uint8_t* mask;
} stage_inputs;

static void free_stage(stage_inputs* s) {
  free(s->frame_ptr);
  free(s->frame_views);
  free(s->H);
  free(s->cosv);
  free(s->sinv);
  free(s->y);
  free(s->w);
  free(s->mask);
  memset(s, 0, sizeof(*s));
}

static int prepare_stage(stage_inputs* s, const float* projections,
                         const float* weights, const float* angles,
                         const float* ptimes, const float* rtimes, int32_t rows,
                         int32_t cols, int32_t num, int32_t recon_num,
                         float vox_wid, float rot_center, float sig_s,
                         float sig_t, float c_s, float c_t, int32_t mxy,
                         int32_t mz, int32_t quad_convex, float hd, float hT,
                         double var_est, const mbir_options* opt, FILE* log) {
  memset(s, 0, sizeof(*s));
  const size_t ns = (size_t)num * rows * cols;
  s->frame_ptr = (int32_t*)calloc(recon_num + 1, sizeof(int32_t));
  s->frame_views = (int32_t*)calloc(num > 0 ? num : 1, sizeof(int32_t));
  if (mbir_view_map(num, ptimes, recon_num, rtimes, s->frame_ptr, s->frame_views) < 0) {
    say(log, "ERROR", "Projection or reconstruction times decrease. Check input files.\n");
    return -1;
  }
  if (mbir_stage_geometry(&s->g, rows, cols, num, recon_num, vox_wid, rot_center, sig_s,
                          sig_t, c_s, c_t, mxy, mz, quad_convex, hd, hT, s->frame_ptr,
                          opt->p_s, opt->p_t)) {
    say(log, "ERROR", "Cannot do reconstruction: mult_xy = %d / mult_z = %d do not divide the detector.\n", mxy, mz);
    return -1;
  }
  s->g.var_est = var_est;
  s->H = (float*)malloc(sizeof(float) * (size_t)num * (RESP_BINS + 1));
  s->cosv = (double*)malloc(sizeof(double) * num);
  s->sinv = (double*)malloc(sizeof(double) * num);
  s->y = (float*)malloc(sizeof(float) * ns);
  s->w = (float*)malloc(sizeof(float) * ns);
  s->mask = (uint8_t*)malloc((size_t)s->g.N_x * s->g.N_y);
  if (!s->H || !s->cosv || !s->sinv || !s->y || !s->w || !s->mask) {
    say(log, "ERROR", "Out of memory.\n");
    return -1;
  }
  mbir_sincos(num, angles, s->cosv, s->sinv);
  mbir_detector_response(&s->g, angles, s->H);
  mbir_sino_to_internal(num, rows, cols, projections, s->y);
  mbir_sino_to_internal(num, rows, cols, weights, s->w);
  for (size_t i = 0; i < ns; ++i) s->w[i] /= var_est; /* XT_Init.c:313-316 */
  mbir_update_mask(&s->g, (double)vox_wid * cols / 2, s->mask);
  return 0;
}

int reconstruct_ex(float** object, float* projections, float* weights,
                   float* proj_angles, float* proj_times, float* recon_times,
                   int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                   int32_t recon_num, float vox_wid, float rot_center,
                   float sig_s, float sig_t, float c_s, float c_t,
                   float convg_thresh, int32_t remove_rings,
                   int32_t quad_convex, float huber_delta, float huber_T,
                   uint8_t restart, FILE* log, const mbir_options* opt_in,
                   int32_t* equits_out, double* cost_out) {
  mbir_options opt;
  int32_t mxy[MBIR_MAX_STAGES], mz[MBIR_MAX_STAGES];
  int32_t nstage, first_stage = 0;
  float* prev = NULL; /* previous stage's object [N_time][N_z][N_y][N_x] */
  int32_t prev_nz = 0, prev_ny = 0, prev_nx = 0;
  double var_est = 1.0; /* VAR_PARAM_INIT, XT_Constants.h:95 */
  float* offset = NULL; /* ring offset [N_r][N_t] carried between stages */
  int rc = -1;
  if (opt_in)
    opt = *opt_in;
  else
    mbir_default_options(&opt);
  if (object) *object = NULL;
  say(log, "INFO", "Reconstructing the data ....\n");
  /* argument checks of XT_Engine.c:101-104 for one process */
  if (!object || !projections || !weights || !proj_angles || !proj_times || !recon_times) {
    say(log, "ERROR", "Null input array.\n");
    return -1;
  }
  if (proj_rows < 4) {
    say(log, "ERROR", "The minimum number of projection rows should be 4.\n");
    return -1;
  }
  if (proj_cols % 2 != 0) {
    say(log, "ERROR", "The number of projection columns should be even.\n");
    return -1;
  }
  if (proj_rows % 2 != 0) {
    say(log, "ERROR", "The number of projection rows should be an even number greater than or equal to 2.\n");
    return -1;
  }
  if (proj_num <= 0 || recon_num <= 0) {
    say(log, "ERROR", "proj_num and recon_num must be positive.\n");
    return -1;
  }
  if (remove_rings < 0 || remove_rings > 3) {
    say(log, "ERROR", "remove_rings must be 0, 1, 2 or 3.\n");
    return -1;
  }
  offset = (float*)calloc((size_t)proj_cols * proj_rows, sizeof(float));
  if (!offset) {
    say(log, "ERROR", "Out of memory.\n");
    return -1;
  }
  nstage = mbir_multires_schedule(proj_rows, proj_cols, mxy, mz);
  if (opt.single_stage) {
    first_stage = nstage - 1;
  } else if (restart == 1) {
    /* the reference resumes from the stage stored in status.bin
     * (XT_Engine.c:143-146), which always holds the first stage run */
    if (read_bin("status", &first_stage, sizeof(int32_t), 1, log)) return -1;
    if (first_stage < 0 || first_stage >= nstage) first_stage = 0;
  }
  const int32_t status_value = first_stage;
  say(log, "INFO", "Number of multi-resolution stages is %d.\n", nstage);
  for (int32_t st = first_stage; st < nstage; ++st) {
    stage_inputs S;
    mbir_hip_ctx* ctx = NULL;
    float* x = NULL;
    int ok = 0;
    say(log, "INFO", "Running multi-resolution stage %d with x-y voxel scale = %d and z voxel scale = %d.\n",
        st, mxy[st], mz[st]);
    if (st > 0 && !opt.single_stage && quad_convex == 0) {
      /* the previous stage's variance estimate (XT_Init.c:286-287) */
      if (opt.write_files) {
        if (read_bin("variance_estimate", &var_est, sizeof(double), 1, log)) goto out;
      }
    }
    if (prepare_stage(&S, projections, weights, proj_angles, proj_times, recon_times,
                      proj_rows, proj_cols, proj_num, recon_num, vox_wid, rot_center, sig_s,
                      sig_t, c_s, c_t, mxy[st], mz[st], quad_convex, huber_delta, huber_T,
                      var_est, &opt, log)) {
      free_stage(&S);
      goto out;
    }
    {
      const mbir_hip_geom* g = &S.g;
      const size_t vol = (size_t)g->N_z * g->N_y * g->N_x;
      x = (float*)calloc(vol * g->N_time, sizeof(float));
      if (!x) {
        say(log, "ERROR", "Out of memory.\n");
        free_stage(&S);
        goto out;
      }
      /* object initialisation (XT_ICD_update.c:540-599): zero at the first
       * stage, else the previous stage's object up-sampled by 2 in x-y
       * (same z scale) or in x-y-z (z scale halved) */
      if (prev && st > first_stage) {
        for (int32_t f = 0; f < g->N_time; ++f) {
          const float* src = prev + (size_t)f * prev_nz * prev_ny * prev_nx;
          float* dst = x + (size_t)f * vol;
          if (mz[st] == mz[st - 1]) {
            for (int32_t z = 0; z < g->N_z; ++z)
              upsample_plane(src + (size_t)z * prev_ny * prev_nx, prev_ny, prev_nx,
                             dst + (size_t)z * g->N_y * g->N_x);
          } else {
            upsample_volume(src, prev_nz, prev_ny, prev_nx, dst);
          }
        }
      } else if (st > 0 && restart == 1 && opt.write_files) {
        /* resumed run: the previous stage's object files */
        const size_t cvol = vol / ((mz[st] == mz[st - 1]) ? 4 : 8);
        const int32_t cz = (mz[st] == mz[st - 1]) ? g->N_z : g->N_z / 2;
        float* c = (float*)malloc(sizeof(float) * cvol);
        char name[64];
        for (int32_t f = 0; f < g->N_time && c; ++f) {
          snprintf(name, sizeof name, "object_time_%d", f);
          if (read_bin(name, c, sizeof(float), cvol, log)) {
            free(c);
            c = NULL;
            break;
          }
          if (mz[st] == mz[st - 1]) {
            for (int32_t z = 0; z < g->N_z; ++z)
              upsample_plane(c + (size_t)z * (g->N_y / 2) * (g->N_x / 2), g->N_y / 2,
                             g->N_x / 2, x + (size_t)f * vol + (size_t)z * g->N_y * g->N_x);
          } else {
            upsample_volume(c, cz, g->N_y / 2, g->N_x / 2, x + (size_t)f * vol);
          }
        }
        if (!c) {
          free(x);
          free_stage(&S);
          goto out;
        }
        free(c);
      }
      /* ring offset (XT_Init.c:246-254, 345-349): updated from the second
       * stage on, starting from zero there and from the previous stage's
       * offset (proj_offset) after it */
      const int update_offset = remove_rings != 0 && st >= 1;
      if (remove_rings != 0 && st >= 2 && st == first_stage && restart == 1 && opt.write_files &&
          read_bin("proj_offset", offset, sizeof(float), (size_t)g->N_r * g->N_t, log))
        goto stage_fail;
      if (mbir_hip_create(&ctx, g, S.H, S.cosv, S.sinv, S.frame_ptr, S.frame_views, S.y, S.w,
                          (remove_rings != 0 && st >= 2) ? offset : NULL, S.mask) ||
          mbir_hip_set_option(ctx, "tile", opt.tile) ||
          (opt.alpha > 0 && mbir_hip_set_option(ctx, "alpha", opt.alpha)) ||
          mbir_hip_set_object(ctx, x) || mbir_hip_build_amatrix(ctx) ||
          mbir_hip_init_error(ctx)) {
        say(log, "ERROR", "GPU stage set-up failed: %s\n", mbir_hip_last_error());
        goto stage_fail;
      }
      double cost[3], last;
      double alpha_now = opt.alpha > 0 ? opt.alpha : g->alpha;
      mbir_guard guard;
      if (mbir_hip_cost(ctx, 0, cost)) goto gpu_fail;
      last = cost[0];
      mbir_guard_init(&guard, cost[0], GUARD_TOL, GUARD_RECOVERIES);
      if (mbir_hip_object_snapshot(ctx, 0)) goto gpu_fail;
      say(log, "INFO", "------------- Iteration 0, Cost = %f------------\n", cost[0]);
      if (opt.write_files && write_bin("cost", &cost[0], sizeof(double), 1, "wb", log)) goto stage_fail;
      int32_t it;
      for (it = 1; it <= opt.max_equits; ++it) {
        double stats[4];
        if (mbir_hip_icd_equits(ctx, 1, it, opt.seed + (uint32_t)st * 7919u, stats)) goto gpu_fail;
        /* the sweeps update e with FP32 atomics; recompute it exactly now
         * and then so rounding drift cannot shift the fixed point */
        if (opt.refresh_every > 0 && it % opt.refresh_every == 0 && mbir_hip_init_error(ctx))
          goto gpu_fail;
        /* ring offset and noise variance (XT_ICD_update.c:983-1050): once
         * per equit here, the reference updates the offset after every
         * half-sweep of its NHICD subsets */
        if (update_offset && mbir_hip_offset_update(ctx, remove_rings)) goto gpu_fail;
        if (st >= 1 && quad_convex == 0 && mbir_hip_variance_update(ctx, &var_est))
          goto gpu_fail;
        /* update magnitude relative to the mean object magnitude of one of
         * the reference's NHICD_Iterations line subsets (XT_ICD_update.c:
         * 963, 1058, 1083-1090) */
        const double pct = stats[1] > 0 ? 100.0 * stats[0] / (stats[1] / NHICD_SUBSETS) : 0.0;
        if (mbir_hip_cost(ctx, 0, cost)) goto gpu_fail;
        say(log, "INFO", "------------- Iteration = %d, Cost = %f, update = %f%% ------------\n",
            it, cost[0], pct);
        if (opt.write_files && write_bin("cost", &cost[0], sizeof(double), 1, "ab", log)) goto stage_fail;
        /* a rise beyond rounding noise: confirmed from a recomputed e, then
         * undone (object restored, alpha lowered) or, once the recoveries
         * are spent, the reference's error (XT_ICD_update.c:1171) */
        int32_t act = mbir_guard_check(&guard, cost[0], 0);
        if (act == MBIR_GUARD_CONFIRM) {
          if (mbir_hip_init_error(ctx) || mbir_hip_cost(ctx, 0, cost)) goto gpu_fail;
          act = mbir_guard_check(&guard, cost[0], 1);
        }
        if (act == MBIR_GUARD_SAVE) {
          if (mbir_hip_object_snapshot(ctx, 0)) goto gpu_fail;
        } else if (act == MBIR_GUARD_RECOVER) {
          const double a2 = mbir_guard_alpha(alpha_now);
          say(log, "WARN", "Cost value increased to %f (lowest %f); restoring that object and "
              "lowering the over-relaxation from %g to %g.\n", cost[0], guard.best, alpha_now, a2);
          alpha_now = a2;
          if (mbir_hip_object_snapshot(ctx, 1) || mbir_hip_init_error(ctx) ||
              mbir_hip_set_option(ctx, "alpha", alpha_now) || mbir_hip_cost(ctx, 0, cost))
            goto gpu_fail;
        } else if (act == MBIR_GUARD_FAIL) {
          say(log, "ERROR", "Cost value increased.\n");
          goto stage_fail;
        }
        last = cost[0];
        if (opt.stop_cost > 0 && cost[0] <= opt.stop_cost) break;
        if (pct < convg_thresh && it > 1) {
          say(log, "INFO", "Convergence criteria is met.\n");
          break;
        }
      }
      if (equits_out) *equits_out = it > opt.max_equits ? opt.max_equits : it;
      if (cost_out) *cost_out = last;
      if (mbir_hip_get_object(ctx, x)) goto gpu_fail;
      if (update_offset && mbir_hip_get_offset(ctx, offset)) goto gpu_fail;
      mbir_hip_destroy(ctx);
      ctx = NULL;
      /* stage files (XT_IOMisc.c:237-269, XT_ICD_update.c:1196-1199,
       * XT_Engine.c:166-168) */
      if (opt.write_files) {
        char name[64];
        for (int32_t f = 0; f < g->N_time; ++f) {
          snprintf(name, sizeof name, "object_time_%d", f);
          if (write_bin(name, x + (size_t)f * vol, sizeof(float), vol, "wb", log)) goto stage_fail;
        }
        ok = !write_bin("proj_offset", offset, sizeof(float), (size_t)g->N_r * g->N_t, "wb", log);
        if (!ok) goto stage_fail;
        if (write_bin("variance_estimate", &var_est, sizeof(double), 1, "wb", log) ||
            write_bin("status", &status_value, sizeof(int32_t), 1, "wb", log))
          goto stage_fail;
      }
      free(prev);
      prev = x;
      x = NULL;
      prev_nz = g->N_z;
      prev_ny = g->N_y;
      prev_nx = g->N_x;
      free_stage(&S);
      say(log, "INFO", "Completed multi-resolution stage %d.\n", st);
      continue;
    }
  gpu_fail:
    say(log, "ERROR", "GPU call failed: %s\n", mbir_hip_last_error());
  stage_fail:
    if (ctx) mbir_hip_destroy(ctx);
    free(x);
    free_stage(&S);
    goto out;
  }
  {
    /* output with halo slices 0 and N_z+1 (XT_Init.c:362-363,
     * XT_Engine.c:170) */
    const size_t plane = (size_t)proj_cols * proj_cols;
    float* obj = (float*)calloc((size_t)recon_num * (proj_rows + 2) * plane, sizeof(float));
    if (!obj || !prev) {
      free(obj);
      say(log, "ERROR", "Out of memory.\n");
      goto out;
    }
    for (int32_t f = 0; f < recon_num; ++f)
      memcpy(obj + ((size_t)f * (proj_rows + 2) + 1) * plane,
             prev + (size_t)f * proj_rows * plane, sizeof(float) * proj_rows * plane);
    *object = obj;
    rc = 0;
    say(log, "INFO", "Exiting MBIR 4D\n");
  }
out:
  free(prev);
  free(offset);
  return rc;
}

int reconstruct(float** object, float* projections, float* weights,
                float* proj_angles, float* proj_times, float* recon_times,
                int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                int32_t recon_num, float vox_wid, float rot_center,
                float sig_s, float sig_t, float c_s, float c_t,
                float convg_thresh, int32_t remove_rings, int32_t quad_convex,
                float huber_delta, float huber_T, uint8_t restart,
                FILE* debug_msg_ptr) {
  return reconstruct_ex(object, projections, weights, proj_angles, proj_times,
                        recon_times, proj_rows, proj_cols, proj_num, recon_num,
                        vox_wid, rot_center, sig_s, sig_t, c_s, c_t,
                        convg_thresh, remove_rings, quad_convex, huber_delta,
                        huber_T, restart, debug_msg_ptr, NULL, NULL, NULL);
}
