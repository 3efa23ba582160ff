/*
 * mbir_oracle.c -- CPU restatement of the reference MBIR_4D hot path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the reference the HIP path is tested against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never does.  All file:line citations are relative to the
 * reference's Source_Code/ directory.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_random2           randlib.c:69-92 known-answer table (first 20 seeds)
 *   orc_view_map          Examples proj_times/recon_times .bin fixtures,
 *                         reference build (oracle/_ref) on random time lists
 *   orc_setup             reference build geometry dump; filter sums to
 *                         mult_xy^3 * delta_recon (closed form)
 *   orc_detector_response bit-exact vs reference build; closed-form area
 *   orc_amatrix_column    bit-exact index/value vs reference build; area
 *                         invariant sum_j A*delta_r ~= delta_xy^2
 *   orc_forward_project   reference build initErrorSinogam; adjointness
 *   orc_qggmrf_weight     finite differences of orc_qggmrf_value; delta=0
 *                         limit q/(sigma^q c)
 *   orc_substitution      reference build; all-equal and p=q closed forms
 *   orc_voxel_thetas      reference build compute_voxel_update_Atten;
 *                         theta2 = ||A_col||^2_W brute force
 *   orc_cost              reference build computeCost; hand-computed 2x2x2
 *   orc_icd_equit         reference ICD cost trajectory is monotone and
 *                         reaches the same converged cost (convexity)
 *   orc_offset_update,    reference build update_Sinogram_Offset /
 *   orc_variance_update   update_variance_parameter; the constraints each
 *                         mode must satisfy; the all-selected variance
 *   orc_phantom_project   the simulator's projection (PCMBIR, whose build
 *                         needs FFTW): identity z response on matched
 *                         slices, mass conservation per view and row,
 *                         linearity (tests/test_simulator_oracle.py)
 * orc_icd_equit_tiled is a schedule study tool with no reference counterpart.
 */
#include "mbir_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PROFILE_RES 1536   /* XT_Constants.h:77 */
#define ORC_BEAM_RES 512       /* XT_Constants.h:78 */
#define ORC_PI 3.14159265358979323846
#define ORC_PI_2 1.57079632679489661923132169163975144
#define ORC_PI_4 0.785398163397448309615660845819875721

/* ------------------------------------------------------------------ RNG */
/* Park-Miller minimal standard generator: s <- 16807 s mod (2^31 - 1),
 * returned as s / (2^31 - 1).  randlib.c:119-141 computes the same product
 * with split 16-bit halves; the 64-bit product here is the textbook form. */
static uint64_t orc_rng_state = 1;

void orc_srandom2(uint32_t seed) { orc_rng_state = seed; }

double orc_random2(void) {
  orc_rng_state = (orc_rng_state * 16807ULL) % 2147483647ULL;
  return (double)orc_rng_state / 2147483647.0;
}

/* ------------------------------------------------------------- set-up */
int32_t orc_view_map(int32_t N_p, const float* proj_times, int32_t N_time,
                     const float* recon_times, int32_t* frame_ptr,
                     int32_t* frame_views) {
  /* XT_Init.c:69-96: frame i takes the views whose time lies in
   * [recon_times[i], recon_times[i+1]); the scan for frame i+1 resumes at
   * the first view whose time reached recon_times[i+1]. */
  int32_t scan_from = 0, n = 0;
  for (int32_t f = 0; f < N_time; ++f) {
    frame_ptr[f] = n;
    for (int32_t v = scan_from; v < N_p; ++v) {
      double tv = proj_times[v];
      if (tv >= recon_times[f] && tv < recon_times[f + 1]) frame_views[n++] = v;
      if (tv >= recon_times[f + 1]) {
        scan_from = v;
        break;
      }
    }
  }
  frame_ptr[N_time] = n;
  return n;
}

static double orc_node_distance(int a, int b, int c, int d) {
  /* XT_Init.c:138-141 */
  return sqrt(pow((double)a - 1.0, 2.0) + pow((double)b - 1.0, 2.0) +
              pow((double)c - 1.0, 2.0) + pow((double)d - 1.0, 2.0));
}

int32_t orc_setup(orc_geom* g, int32_t proj_rows, int32_t proj_cols,
                  int32_t proj_num, int32_t recon_num, float vox_wid,
                  float rot_center, float sig_s, float sig_t, float c_s,
                  float c_t, int32_t mult_xy, int32_t mult_z,
                  int32_t quad_convex, float huber_delta, float huber_T,
                  const int32_t* frame_ptr, double p_s, double p_t) {
  memset(g, 0, sizeof(*g));
  const double vw = vox_wid, rc = rot_center;
  const double length_r = vw * proj_cols;  /* XT_Init.c:223-224 */
  const double length_t = vw * proj_rows;  /* single rank: no split */
  g->N_p = proj_num;
  g->N_r = proj_cols;
  g->N_t = proj_rows;
  g->N_time = recon_num;
  g->mult_xy = mult_xy;
  g->mult_z = mult_z;
  g->W = 2 * mult_xy + 1;
  /* XT_Init.c:318-323 */
  g->delta_r = length_r / g->N_r;
  g->delta_t = length_t / g->N_t;
  g->R0 = -rc * g->delta_r;
  g->RMax = (g->N_r - rc) * g->delta_r;
  g->T0 = -length_t / 2.0;
  g->TMax = length_t / 2.0;
  /* XT_Init.c:329-334, 358-361 */
  g->N_x = (int32_t)(g->N_r / (double)mult_xy);
  g->N_y = g->N_x;
  g->N_z = (int32_t)(g->N_t / (double)mult_z);
  g->delta_xy = mult_xy * g->delta_r;
  g->delta_z = mult_z * g->delta_t;
  g->x0 = g->R0;
  g->z0 = g->T0;
  g->y0 = -length_r / 2.0;
  /* XT_Init.c:372-373 */
  g->OffsetR = (g->delta_xy / sqrt(2.0) + g->delta_r / 2.0) / ORC_RESPONSE_BINS;
  g->OffsetT = ((g->delta_z / 2) + g->delta_t / 2) / ORC_RESPONSE_BINS;
  g->radius_obj = vw * proj_cols / 2;  /* XT_Init.c:243 */
  /* XT_Init.c:117-126: mean views per frame */
  double nviews = 0;
  for (int32_t f = 0; f < recon_num; ++f) nviews += frame_ptr[f + 1] - frame_ptr[f];
  g->delta_recon = nviews / recon_num;
  /* XT_Init.c:418-421 (one MRF_P in the reference; p_s/p_t here) */
  g->q = 2.0;
  g->p_s = p_s;
  g->p_t = p_t;
  g->c_s = c_s;
  g->c_t = c_t;
  const double ss = (double)sig_s * (double)mult_xy;
  const double st = (double)sig_t * g->delta_recon;
  g->sig_s_q = pow(ss, g->q);
  g->sig_s_qp = pow(ss, g->q - p_s);
  g->sig_t_q = pow(st, g->q);
  g->sig_t_qp = pow(st, g->q - p_t);
  /* XT_Init.c:144-173: 1/distance weights, normalised with the two
   * temporal neighbours included, scaled by mult_xy^3 * delta_recon. */
  const double scale = (double)mult_xy * mult_xy * mult_xy * g->delta_recon;
  double ft = 1.0 / orc_node_distance(0, 1, 1, 1);
  double sum = 2.0 * ft;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double v = 0;
        if (a != 1 || b != 1 || c != 1) {
          v = 1.0 / orc_node_distance(1, a, b, c);
          sum = sum + v;
        }
        g->filt[(a * 3 + b) * 3 + c] = v;
      }
  for (int k = 0; k < 27; ++k) g->filt[k] = scale * g->filt[k] / sum;
  g->filt_t = scale * ft / sum;
  g->alpha = 1.5; /* XT_Constants.h:87 */
  if (quad_convex) { /* XT_Init.c:259-268, XT_Constants.h:97-98 */
    g->huber_T = 100000;
    g->huber_delta = 1;
  } else {
    g->huber_T = huber_T;
    g->huber_delta = huber_delta;
  }
  g->var_est = 1.0;
  /* XT_Init.c:427-428 */
  if (g->N_t % mult_z != 0 || g->N_r % mult_xy != 0) return -1;
  return 0;
}

void orc_sincos(int32_t N_p, const float* angles, double* cosv, double* sinv) {
  for (int32_t i = 0; i < N_p; ++i) { /* XT_Init.c:200-204 */
    cosv[i] = cos(angles[i]);
    sinv[i] = sin(angles[i]);
  }
}

/* ---------------------------------------------- detector response (a2) */
void orc_detector_response(const orc_geom* g, const float* angles, float* H) {
  const int32_t NB = ORC_RESPONSE_BINS + 1;
  float* prof = (float*)malloc(sizeof(float) * ORC_PROFILE_RES);
  double beam[ORC_BEAM_RES];
  /* Hamming beam over one detector pitch, normalised to unit area
   * (XT_Profile.c:108-124). */
  double bsum = 0;
  for (int i = 0; i < ORC_BEAM_RES; ++i) {
    beam[i] = 0.54 - 0.46 * cos((2 * ORC_PI / ORC_BEAM_RES) * i);
    bsum = bsum + beam[i];
  }
  for (int i = 0; i < ORC_BEAM_RES; ++i) beam[i] /= bsum;

  const double bw = g->delta_r; /* BeamWidth, XT_Init.c:361 */
  const double r_first = -bw / 2;
  const double step = bw / ORC_BEAM_RES;
  const double to_index = (ORC_PROFILE_RES) / (2 * g->delta_xy);

  for (int32_t v = 0; v < g->N_p; ++v) {
    /* Trapezoidal line integral of a square voxel at this view, sampled on
     * [0, 2) in units of delta_xy (XT_Profile.c:47-104). */
    double ang = angles[v];
    while (ang > ORC_PI_2) ang -= ORC_PI_2;
    while (ang < 0) ang += ORC_PI_2;
    double peak = (ang <= ORC_PI_4) ? g->delta_xy / cos(ang)
                                    : g->delta_xy / cos(ORC_PI_2 - ang);
    double c45 = cos(ORC_PI_4);
    double d_outer = c45 * cos((ORC_PI_4 - ang));
    double d_inner = c45 * fabs((cos((ORC_PI_4 + ang))));
    double lo_corner = 1 - d_outer, lo_flat = 1 - d_inner;
    double hi_flat = 1 + d_inner, hi_corner = 1 + d_outer;
    for (int j = 0; j < ORC_PROFILE_RES; ++j) {
      double t = 2.0 * j / ORC_PROFILE_RES;
      double val;
      if (t <= lo_corner || t >= hi_corner)
        val = 0;
      else if (t > hi_flat)
        val = peak * (hi_corner - t) / (hi_corner - hi_flat);
      else if (t >= lo_flat)
        val = peak;
      else
        val = peak * (t - lo_corner) / (lo_flat - lo_corner);
      prof[j] = (float)val;
    }
    /* Convolve with the beam at 64 offsets of OffsetR (XT_Profile.c:
     * 143-162); the 65th entry is zero. */
    for (int i = 0; i < ORC_RESPONSE_BINS; ++i) {
      double centre = i * g->OffsetR;
      double lo = centre - g->delta_xy;
      double acc = 0;
      for (int p = 0; p < ORC_BEAM_RES; ++p) {
        double r = r_first + p * step;
        int32_t pi = (int32_t)floor((r - lo) * to_index);
        if (pi >= 0 && pi <= ORC_PROFILE_RES - 1) acc += (prof[pi] * beam[p]);
      }
      H[v * NB + i] = (float)acc;
    }
    H[v * NB + ORC_RESPONSE_BINS] = 0.0f;
  }
  free(prof);
}

/* --------------------------------------------------- A column (a1) */
int32_t orc_amatrix_column(const orc_geom* g, const double* cosv,
                           const double* sinv, const float* H, int32_t row,
                           int32_t col, int32_t view, int32_t* idx,
                           double* val) {
  const int32_t NB = ORC_RESPONSE_BINS + 1;
  double xc = g->x0 + ((double)col + 0.5) * g->delta_xy;
  double yc = g->y0 + ((double)row + 0.5) * g->delta_xy;
  double r = xc * cosv[view] - yc * sinv[view];
  double rlo = r - g->delta_xy, rhi = r + g->delta_xy;
  if (rhi < g->R0 || rlo > g->RMax) return 0;
  int32_t jlo = (int32_t)(floor(((rlo - g->R0) / g->delta_r)));
  int32_t jhi = (int32_t)(floor((rhi - g->R0) / g->delta_r));
  if (jhi >= g->N_r) jhi = g->N_r - 1;
  if (jlo < 0) jlo = 0;
  int32_t n = 0;
  const float* h = H + (size_t)view * NB;
  for (int32_t j = jlo; j <= jhi; ++j) {
    double centre = (g->R0 + (((double)j) + 0.5) * (g->delta_r));
    double d = fabs(r - centre);
    int32_t k = (int32_t)(floor((d / g->OffsetR)));
    if (k >= 0 && k < ORC_RESPONSE_BINS) {
      double w_hi = d - k * g->OffsetR;
      double w_lo = g->OffsetR - w_hi;
      double f = (w_lo * h[k] + w_hi * h[k + 1]) / g->OffsetR;
      if (f > 0) {
        idx[n] = j;
        val[n] = f;
        ++n;
      }
    }
  }
  return n;
}

int32_t orc_phantom_project(const float* mag, const float* phase, int32_t n_xy,
                            int32_t n_z, int32_t n_p, const float* angles,
                            int32_t proj_rows, int32_t proj_cols, double vox_wid,
                            double rot_center, double wavenumber, double* out_r,
                            double* out_i) {
  if (proj_cols <= 0 || n_xy % proj_cols != 0) return -1;
  orc_geom g;
  memset(&g, 0, sizeof g);
  /* initPhantomStructures: the detector sampled at the phantom's voxels */
  const double Length_R = vox_wid * proj_cols, Length_T = vox_wid * proj_rows;
  const double rc = rot_center * (double)(n_xy / proj_cols);
  g.N_r = n_xy;
  g.N_t = n_z;
  g.N_p = n_p;
  g.delta_r = Length_R / g.N_r;
  g.delta_t = Length_T / g.N_t;
  g.R0 = -rc * g.delta_r;
  g.RMax = (g.N_r - rc) * g.delta_r;
  g.T0 = -Length_T / 2.0;
  g.TMax = Length_T / 2.0;
  g.N_x = g.N_y = n_xy;
  g.N_z = n_z;
  g.delta_xy = g.delta_r;
  g.delta_z = g.delta_t;
  g.x0 = g.R0;
  g.z0 = g.T0;
  g.y0 = -Length_R / 2.0;
  g.OffsetR = (g.delta_xy / sqrt(2.0) + g.delta_r / 2.0) / ORC_RESPONSE_BINS;
  g.OffsetT = ((g.delta_z / 2) + g.delta_t / 2) / ORC_RESPONSE_BINS;
  g.W = 3;
  /* z-line response table H_t (XT_Profile.c:182-227), float like the
   * reference's Real_arr_t */
  float Ht[ORC_RESPONSE_BINS + 1];
  for (int32_t i = 0; i < ORC_RESPONSE_BINS; ++i) {
    const double c = i * g.OffsetT;
    double h;
    if (g.delta_z >= g.delta_t)
      h = (c <= g.delta_z / 2 - g.delta_t / 2) ? g.delta_t : -c + g.delta_z / 2 + g.delta_t / 2;
    else
      h = (c <= g.delta_t / 2 - g.delta_z / 2) ? g.delta_z : -c + g.delta_z / 2 + g.delta_t / 2;
    if (h < 0) h = 0;
    Ht[i] = (float)(h / g.delta_t);
  }
  Ht[ORC_RESPONSE_BINS] = 0.f;
  /* voxel line response per slice (XT_Profile.c:229-290): up to 3 rows */
  int32_t* lt = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)n_z);
  double* lv = (double*)malloc(sizeof(double) * 3 * (size_t)n_z);
  int32_t* ln = (int32_t*)calloc((size_t)n_z, sizeof(int32_t));
  for (int32_t i = 0; i < n_z; ++i) {
    const double t = ((double)i + 0.5) * g.delta_z + g.z0;
    const double tmin = (t - g.delta_z / 2) > g.T0 ? t - g.delta_z / 2 : g.T0;
    const double tmax = (t + g.delta_z / 2) <= g.TMax ? t + g.delta_z / 2 : g.TMax;
    int32_t lo = (int32_t)floor((tmin - g.T0) / g.delta_t);
    int32_t hi = (int32_t)floor((tmax - g.T0) / g.delta_t);
    if (lo < 0) lo = 0;
    if (hi >= g.N_t) hi = g.N_t - 1;
    for (int32_t it = lo; it <= hi; ++it) {
      const double ct = ((double)it + 0.5) * g.delta_t + g.T0;
      const double d = fabs(ct - t);
      const int32_t k = (int32_t)floor(d / g.OffsetT);
      double v = 0;
      if (k < ORC_RESPONSE_BINS) {
        const double w3 = d - (double)k * g.OffsetT;
        const double w4 = ((double)k + 1) * g.OffsetT - d;
        v = (w4 / g.OffsetT) * Ht[k] + (w3 / g.OffsetT) * Ht[k + 1];
      }
      if (v != 0 && ln[i] < 3) {
        lt[3 * i + ln[i]] = it;
        lv[3 * i + ln[i]] = v;
        ++ln[i];
      }
    }
  }
  float* H = (float*)malloc(sizeof(float) * (size_t)n_p * (ORC_RESPONSE_BINS + 1));
  double* cs = (double*)malloc(sizeof(double) * (size_t)n_p);
  double* sn = (double*)malloc(sizeof(double) * (size_t)n_p);
  orc_detector_response(&g, angles, H);
  orc_sincos(n_p, angles, cs, sn);
  const size_t per = (size_t)n_z * n_xy;
  memset(out_r, 0, sizeof(double) * per * n_p);
  memset(out_i, 0, sizeof(double) * per * n_p);
#pragma omp parallel for schedule(dynamic)
  for (int32_t p = 0; p < n_p; ++p) {
    int32_t idx[64];
    double val[64];
    double* orp = out_r + (size_t)p * per;
    double* oip = out_i + (size_t)p * per;
    for (int32_t y = 0; y < n_xy; ++y)
      for (int32_t x = 0; x < n_xy; ++x) {
        const int32_t n = orc_amatrix_column(&g, cs, sn, H, y, x, p, idx, val);
        for (int32_t z = 0; z < n_z; ++z) {
          const double m = mag[((size_t)z * n_xy + y) * n_xy + x];
          const double ph = phase[((size_t)z * n_xy + y) * n_xy + x];
          if (m == 0 && ph == 0) continue;
          for (int32_t a = 0; a < n; ++a) {
            const double av = val[a] * wavenumber;
            for (int32_t q = 0; q < ln[z]; ++q) {
              const size_t o = (size_t)lt[3 * z + q] * n_xy + idx[a];
              orp[o] += m * av * lv[3 * z + q];
              oip[o] += ph * av * lv[3 * z + q];
            }
          }
        }
      }
  }
  free(lt);
  free(lv);
  free(ln);
  free(H);
  free(cs);
  free(sn);
  return 0;
}

void orc_amatrix_all(const orc_geom* g, const double* cosv, const double* sinv,
                     const float* H, int32_t* first, double* val) {
  const int32_t W = g->W;
#pragma omp parallel
  {
    int32_t idx[64];
    double v[64];
#pragma omp for schedule(static)
    for (int32_t c = 0; c < g->N_x * g->N_y; ++c) {
      int32_t yy = c / g->N_x, xx = c % g->N_x;
      for (int32_t p = 0; p < g->N_p; ++p) {
        size_t o = (size_t)c * g->N_p + p;
        double* dst = val + o * W;
        for (int k = 0; k < W; ++k) dst[k] = 0;
        int32_t n = orc_amatrix_column(g, cosv, sinv, H, yy, xx, p, idx, v);
        if (n == 0) {
          first[o] = -1;
          continue;
        }
        /* window starts at the clamped lower index; kept bins are
         * contiguous around the projected centre */
        double rlo = g->x0 + ((double)xx + 0.5) * g->delta_xy;
        double ylo = g->y0 + ((double)yy + 0.5) * g->delta_xy;
        double r = rlo * cosv[p] - ylo * sinv[p] - g->delta_xy;
        int32_t j0 = (int32_t)floor((r - g->R0) / g->delta_r);
        if (j0 < 0) j0 = 0;
        first[o] = j0;
        for (int32_t m = 0; m < n; ++m) dst[idx[m] - j0] = v[m];
      }
    }
  }
}

/* ------------------------------------------- forward projection (a3) */
void orc_forward_project(const orc_geom* g, const double* cosv,
                         const double* sinv, const float* H,
                         const int32_t* frame_ptr, const int32_t* frame_views,
                         const double* x, double* ax) {
  const int32_t Nr = g->N_r, Nt = g->N_t, zo = g->mult_z;
  memset(ax, 0, sizeof(double) * (size_t)g->N_p * Nr * Nt);
  /* frames own disjoint views, so frames run in parallel without races
   * (XT_ICD_update.c:665-682) */
#pragma omp parallel for schedule(dynamic)
  for (int32_t f = 0; f < g->N_time; ++f) {
    int32_t idx[64];
    double v[64];
    for (int32_t yy = 0; yy < g->N_y; ++yy)
      for (int32_t xx = 0; xx < g->N_x; ++xx)
        for (int32_t s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
          int32_t p = frame_views[s];
          int32_t n = orc_amatrix_column(g, cosv, sinv, H, yy, xx, p, idx, v);
          for (int32_t z = 0; z < g->N_z; ++z) {
            double pix = x[(((size_t)f * g->N_z + z) * g->N_y + yy) * g->N_x + xx];
            for (int32_t m = 0; m < n; ++m)
              for (int32_t k = 0; k < zo; ++k)
                ax[((size_t)p * Nr + idx[m]) * Nt + z * zo + k] += pix * v[m];
          }
        }
  }
}

/* --------------------------------------------------- qGGMRF (a5) */
double orc_qggmrf_value(const orc_geom* g, double delta, int spatial) {
  /* XT_ICD_update.c:82-90 */
  double sq = spatial ? g->sig_s_q : g->sig_t_q;
  double sqp = spatial ? g->sig_s_qp : g->sig_t_qp;
  double c = spatial ? g->c_s : g->c_t;
  double p = spatial ? g->p_s : g->p_t;
  double ad = fabs(delta);
  return ((pow(ad, g->q) / sq) / (c + pow(ad, g->q - p) / sqp));
}

double orc_qggmrf_weight(const orc_geom* g, double delta, int spatial) {
  double sq = spatial ? g->sig_s_q : g->sig_t_q;
  double sqp = spatial ? g->sig_s_qp : g->sig_t_qp;
  double c = spatial ? g->c_s : g->c_t;
  double p = spatial ? g->p_s : g->p_t;
  if (delta == 0) return g->q / (sq * c); /* rho''(0), XT_ICD_update.c:120-128 */
  /* rho'(delta) (XT_ICD_update.c:92-118), divided by delta */
  double ratio = pow(fabs(delta), g->q - p) / (sqp);
  double mag = pow(fabs(delta), g->q - 1);
  double den = c + ratio;
  double shape = (g->q - ((g->q - p) * ratio) / (den));
  double deriv = (delta < 0) ? ((-1 * mag / (den * sq)) * shape)
                             : ((mag / (den * sq)) * shape);
  return deriv / delta;
}

double orc_substitution(const orc_geom* g, double V, double theta1,
                        double theta2, const double* nb, const uint8_t* nbf,
                        const double* tn, const uint8_t* tf) {
  /* XT_ICD_update.c:133-182: one step of the quadratic surrogate of the
   * qGGMRF prior around V, plus the data term's theta1/theta2 parabola,
   * then over-relaxation by alpha. */
  double num = 0, den = 0;
  for (int k = 0; k < 27; ++k) {
    if (k == 13 || !nbf[k]) continue;
    double wgt = g->filt[k] * orc_qggmrf_weight(g, V - nb[k], 1);
    num += wgt * nb[k];
    den += wgt;
  }
  for (int k = 0; k < 2; ++k) {
    if (!tf[k]) continue;
    double wgt = g->filt_t * orc_qggmrf_weight(g, V - tn[k], 0);
    num += wgt * tn[k];
    den += wgt;
  }
  double u = (num + (theta2 * V) - theta1) / (den + theta2);
  return V + g->alpha * (u - V);
}

/* --------------------------------------------- voxel thetas (a4) */
static void orc_thetas_cols(const orc_geom* g, int32_t f, int32_t z,
                            const int32_t* frame_ptr, const int32_t* frame_views,
                            const int32_t* cnt, const int32_t* cidx,
                            const double* cval, const double* e,
                            const double* w, double* th1, double* th2) {
  /* XT_VoxUpdate.c:56-85: sums over the frame's views, the kept bins and
   * the z_overlap detector rows of the voxel's slice; entries flagged as
   * outliers (|e sqrt(w)| >= T) use the generalized-Huber weight
   * T delta sqrt(w) / |e|. */
  const int32_t Nr = g->N_r, Nt = g->N_t, zo = g->mult_z, cap = g->W;
  double t1 = 0, t2 = 0;
  for (int32_t s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
    int32_t p = frame_views[s];
    int32_t si = s - frame_ptr[f];
    for (int32_t m = 0; m < cnt[si]; ++m) {
      int32_t r = cidx[si * cap + m];
      double a = (cval[si * cap + m] * g->delta_t);
      for (int32_t k = 0; k < zo; ++k) {
        size_t o = ((size_t)p * Nr + r) * Nt + z * zo + k;
        double ev = e[o], wv = w[o];
        if (fabs(ev * sqrt(wv)) < g->huber_T) {
          t2 += (a * a * wv);
          t1 += -(ev * a * wv);
        } else {
          double hw = g->huber_T * g->huber_delta * sqrt(wv) / fabs(ev);
          t2 += (a * a * hw);
          t1 += -(ev * a * hw);
        }
      }
    }
  }
  *th1 = t1;
  *th2 = t2;
}

void orc_voxel_thetas(const orc_geom* g, const double* cosv,
                      const double* sinv, const float* H,
                      const int32_t* frame_ptr, const int32_t* frame_views,
                      const double* e, const double* w, int32_t t, int32_t z,
                      int32_t yy, int32_t xx, double* theta1, double* theta2) {
  int32_t nv = frame_ptr[t + 1] - frame_ptr[t];
  int32_t cap = g->W;
  int32_t* cnt = (int32_t*)calloc(nv + 1, sizeof(int32_t));
  int32_t* cidx = (int32_t*)calloc((size_t)(nv + 1) * cap, sizeof(int32_t));
  double* cval = (double*)calloc((size_t)(nv + 1) * cap, sizeof(double));
  for (int32_t s = 0; s < nv; ++s)
    cnt[s] = orc_amatrix_column(g, cosv, sinv, H, yy, xx,
                                frame_views[frame_ptr[t] + s], cidx + s * cap,
                                cval + s * cap);
  orc_thetas_cols(g, t, z, frame_ptr, frame_views, cnt, cidx, cval, e, w,
                  theta1, theta2);
  free(cnt);
  free(cidx);
  free(cval);
}

/* ------------------------------------------------------- cost (a8) */
double orc_cost(const orc_geom* g, const double* e, const double* w,
                const double* x, double* out) {
  const size_t ns = (size_t)g->N_p * g->N_r * g->N_t;
  const double T = g->huber_T, dl = g->huber_delta;
  double data = 0;
#pragma omp parallel for reduction(+ : data) schedule(static)
  for (size_t i = 0; i < ns; ++i) {
    double s = e[i] * sqrt(w[i]);
    if (fabs(s) < T)
      data += s * s;
    else
      data += 2.0 * dl * T * fabs(s) + T * T * (1.0 - 2.0 * dl);
  }
  data /= 2.0;
  data += (double)g->N_p * (double)g->N_r * (double)g->N_t * log(g->var_est) / 2;
  /* Prior: each neighbour pair once -- 4 in-plane forward neighbours, 9 at
   * z+1, and t+1 (XT_ICD_update.c:205-298). */
  static const int fwd[13][3] = {{0, 0, 1},  {0, 1, -1}, {0, 1, 0},
                                 {0, 1, 1},  {1, -1, 0}, {1, 0, 0},
                                 {1, 1, 0},  {1, -1, -1}, {1, -1, 1},
                                 {1, 0, -1}, {1, 1, -1}, {1, 1, 1},
                                 {1, 0, 1}};
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z, Nf = g->N_time;
  double prior = 0;
#pragma omp parallel for reduction(+ : prior) schedule(static)
  for (int32_t fz = 0; fz < Nf * Nz; ++fz) {
    int32_t f = fz / Nz, z = fz % Nz;
    for (int32_t yy = 0; yy < Ny; ++yy)
      for (int32_t xx = 0; xx < Nx; ++xx) {
        size_t o = (((size_t)f * Nz + z) * Ny + yy) * Nx + xx;
        double v = x[o];
        for (int n = 0; n < 13; ++n) {
          int32_t z2 = z + fwd[n][0], y2 = yy + fwd[n][1], x2 = xx + fwd[n][2];
          if (z2 >= Nz || y2 < 0 || y2 >= Ny || x2 < 0 || x2 >= Nx) continue;
          double nbv = x[(((size_t)f * Nz + z2) * Ny + y2) * Nx + x2];
          int k = ((fwd[n][0] + 1) * 3 + (fwd[n][1] + 1)) * 3 + fwd[n][2] + 1;
          prior += g->filt[k] * orc_qggmrf_value(g, v - nbv, 1);
        }
        if (f + 1 < Nf) {
          double nbv = x[o + (size_t)Nz * Ny * Nx];
          prior += g->filt_t * orc_qggmrf_value(g, v - nbv, 0);
        }
      }
  }
  if (out) {
    out[0] = data;
    out[1] = prior;
  }
  return data + prior;
}

void orc_mask(const orc_geom* g, uint8_t* mask) {
  for (int32_t yy = 0; yy < g->N_y; ++yy)
    for (int32_t xx = 0; xx < g->N_x; ++xx) {
      double xc = g->x0 + ((double)xx + 0.5) * g->delta_xy;
      double yc = g->y0 + ((double)yy + 0.5) * g->delta_xy;
      mask[yy * g->N_x + xx] =
          (xc * xc + yc * yc < g->radius_obj * g->radius_obj) ? 1 : 0;
    }
}

/* ---------------------------------------------- one voxel update */
/* Gathers the neighbourhood of (f, z, yy, xx), applies zero skipping and
 * the substitution, writes x and updates e along the column
 * (XT_VoxUpdate.c:92-116, 165-270).  Returns 1 if skipped. */
static int orc_update_voxel(const orc_geom* g, int32_t f, int32_t z,
                            int32_t yy, int32_t xx, int32_t iter,
                            const int32_t* frame_ptr,
                            const int32_t* frame_views, const int32_t* cnt,
                            const int32_t* cidx, const double* cval, double* x,
                            double* e, const double* w, double* dabs,
                            double* xabs) {
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z;
  const size_t vol = (size_t)Nz * Ny * Nx;
  double nb[27];
  uint8_t nbf[27];
  double tn[2];
  uint8_t tf[2];
  size_t o = (((size_t)f * Nz + z) * Ny + yy) * Nx + xx;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int k = ((dz + 1) * 3 + dy + 1) * 3 + dx + 1;
        int32_t z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
        if (z2 >= 0 && z2 < Nz && y2 >= 0 && y2 < Ny && x2 >= 0 && x2 < Nx) {
          nb[k] = x[(((size_t)f * Nz + z2) * Ny + y2) * Nx + x2];
          nbf[k] = 1;
        } else {
          nb[k] = 0;
          nbf[k] = 0;
        }
      }
  nb[13] = 0;
  tf[0] = f > 0;
  tn[0] = tf[0] ? x[o - vol] : 0;
  tf[1] = f + 1 < g->N_time;
  tn[1] = tf[1] ? x[o + vol] : 0;
  double V = x[o];
  /* zero skipping (XT_VoxUpdate.c:238-258) */
  if (V == 0.0 && iter > 1) {
    int any = (tn[0] > 0.0 || tn[1] > 0.0);
    for (int k = 0; k < 27 && !any; ++k)
      if (nb[k] > 0.0) any = 1;
    if (!any) return 1;
  }
  double t1, t2;
  orc_thetas_cols(g, f, z, frame_ptr, frame_views, cnt, cidx, cval, e, w, &t1,
                  &t2);
  double xn = orc_substitution(g, V, t1, t2, nb, nbf, tn, tf);
  x[o] = xn;
  const int32_t Nr = g->N_r, Nt = g->N_t, zo = g->mult_z, cap = g->W;
  for (int32_t s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
    int32_t p = frame_views[s], si = s - frame_ptr[f];
    for (int32_t m = 0; m < cnt[si]; ++m) {
      double a = cval[si * cap + m] * g->delta_t;
      int32_t r = cidx[si * cap + m];
      for (int32_t k = 0; k < zo; ++k)
        e[((size_t)p * Nr + r) * Nt + z * zo + k] -= a * (xn - V);
    }
  }
  *dabs += fabs(xn - V);
  *xabs += fabs(xn);
  return 0;
}

static void orc_line_columns(const orc_geom* g, const double* cosv,
                             const double* sinv, const float* H, int32_t f,
                             int32_t yy, int32_t xx, const int32_t* frame_ptr,
                             const int32_t* frame_views, int32_t* cnt,
                             int32_t* cidx, double* cval) {
  for (int32_t s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
    int32_t si = s - frame_ptr[f];
    cnt[si] = orc_amatrix_column(g, cosv, sinv, H, yy, xx, frame_views[s],
                                 cidx + si * g->W, cval + si * g->W);
  }
}

/* ------------------------------------------------ ICD equit (a6/a7) */
void orc_icd_equit(const orc_geom* g, const double* cosv, const double* sinv,
                   const float* H, const int32_t* frame_ptr,
                   const int32_t* frame_views, const uint8_t* mask,
                   int32_t num_z_blocks, int32_t K, int32_t iter, double* x,
                   double* e, const double* w, double* stats) {
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z, Nf = g->N_time;
  const int32_t ncol = Nx * Ny, nb = num_z_blocks;
  /* Random line order per (frame, block) (XT_ICD_update.c:499-531): draw
   * without replacement from all columns, keep the masked ones. */
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)Nf * nb * ncol);
  int32_t* count = (int32_t*)calloc((size_t)Nf * nb, sizeof(int32_t));
  int32_t* pool = (int32_t*)malloc(sizeof(int32_t) * ncol);
  const int32_t zper = Nz / nb;
  for (int32_t f = 0; f < Nf; ++f)
    for (int32_t b = 0; b < nb; ++b) {
      int32_t left = ncol;
      for (int32_t i = 0; i < ncol; ++i) pool[i] = i;
      int32_t* dst = order + ((size_t)f * nb + b) * ncol;
      for (int32_t i = 0; i < ncol; ++i) {
        int32_t pick = (int32_t)floor(orc_random2() * left);
        if (pick == left) pick = left - 1;
        int32_t c = pool[pick];
        if (zper > 0 && mask[c]) dst[count[f * nb + b]++] = c;
        pool[pick] = pool[left - 1];
        --left;
      }
    }
  free(pool);
  double dsum = 0, xsum = 0, skipped = 0;
  for (int32_t j = 0; j < K; ++j) {
    for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for collapse(2) schedule(dynamic) reduction(+ : dsum, xsum, skipped)
      for (int32_t f = 0; f < Nf; ++f)
        for (int32_t bb = 0; bb < nb; bb += 2) {
          /* XT_ICD_update.c:966-974, 989-997 */
          int32_t b = ((f % 2 == 0) == (phase == 0)) ? bb : bb + 1;
          if (b >= nb) continue;
          int32_t z0 = b * zper;
          int32_t z1 = (b >= nb - 1) ? Nz - 1 : (b + 1) * zper - 1;
          int32_t n = count[f * nb + b];
          int32_t lo = j * (n / K);
          int32_t hi = (j == K - 1) ? n - 1 : (j + 1) * (n / K) - 1;
          int32_t nv = frame_ptr[f + 1] - frame_ptr[f];
          int32_t* cnt = (int32_t*)calloc(nv + 1, sizeof(int32_t));
          int32_t* cidx = (int32_t*)calloc((size_t)(nv + 1) * g->W, sizeof(int32_t));
          double* cval = (double*)calloc((size_t)(nv + 1) * g->W, sizeof(double));
          const int32_t* ord = order + ((size_t)f * nb + b) * ncol;
          for (int32_t i = lo; i <= hi; ++i) {
            int32_t c = ord[i], yy = c / Nx, xx = c % Nx;
            orc_line_columns(g, cosv, sinv, H, f, yy, xx, frame_ptr,
                             frame_views, cnt, cidx, cval);
            for (int32_t z = z0; z <= z1; ++z)
              skipped += orc_update_voxel(g, f, z, yy, xx, iter, frame_ptr,
                                          frame_views, cnt, cidx, cval, x, e,
                                          w, &dsum, &xsum);
          }
          free(cnt);
          free(cidx);
          free(cval);
        }
    }
  }
  free(order);
  free(count);
  if (stats) {
    stats[0] = dsum;
    stats[1] = xsum;
    stats[2] = skipped;
  }
}

/* ---------------------------------------- tiled schedule study tool */
static uint32_t orc_hash(uint32_t a) {
  a ^= a >> 16;
  a *= 0x7feb352dU;
  a ^= a >> 15;
  a *= 0x846ca68bU;
  a ^= a >> 16;
  return a;
}

void orc_icd_equit_tiled(const orc_geom* g, const double* cosv,
                         const double* sinv, const float* H,
                         const int32_t* frame_ptr, const int32_t* frame_views,
                         const uint8_t* mask, int32_t T, double alpha,
                         uint32_t seed, int32_t iter, double* x, double* e,
                         const double* w) {
  orc_icd_equit_tiled_sync(g, cosv, sinv, H, frame_ptr, frame_views, mask, T,
                           alpha, seed, iter, 1, x, e, w);
}

void orc_icd_equit_tiled_sync(const orc_geom* g, const double* cosv,
                              const double* sinv, const float* H,
                              const int32_t* frame_ptr,
                              const int32_t* frame_views, const uint8_t* mask,
                              int32_t T, double alpha, uint32_t seed,
                              int32_t iter, int32_t m, double* x, double* e,
                              const double* w) {
  orc_geom gg = *g;
  gg.alpha = alpha;
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z, Nf = g->N_time;
  const int32_t TX = (Nx + T - 1) / T, TY = (Ny + T - 1) / T;
  const size_t vol = (size_t)Nz * Ny * Nx;
  const size_t ns = (size_t)g->N_p * g->N_r * g->N_t;
  double* esnap = (double*)malloc(sizeof(double) * ns);
  int32_t* lists = (int32_t*)malloc(sizeof(int32_t) * (size_t)TX * TY * Nf * T * T);
  int32_t* lens = (int32_t*)calloc((size_t)TX * TY * Nf, sizeof(int32_t));
  /* per (tile, frame) list of masked columns in a seeded random order */
  for (int32_t ty = 0; ty < TY; ++ty)
    for (int32_t tx = 0; tx < TX; ++tx)
      for (int32_t f = 0; f < Nf; ++f) {
        size_t id = ((size_t)ty * TX + tx) * Nf + f;
        int32_t* L = lists + id * T * T;
        int32_t n = 0;
        for (int32_t yy = ty * T; yy < ty * T + T && yy < Ny; ++yy)
          for (int32_t xx = tx * T; xx < tx * T + T && xx < Nx; ++xx)
            if (mask[yy * Nx + xx]) L[n++] = yy * Nx + xx;
        uint32_t h = orc_hash(seed * 2654435761U + (uint32_t)id);
        for (int32_t i = n - 1; i > 0; --i) {
          h = orc_hash(h + (uint32_t)i);
          int32_t j = (int32_t)(h % (uint32_t)(i + 1));
          int32_t tmp = L[i];
          L[i] = L[j];
          L[j] = tmp;
        }
        lens[id] = n;
      }
  for (int color = 0; color < 4; ++color)
    for (int fpar = 0; fpar < 2; ++fpar) {
      int32_t maxlen = 0;
      for (int32_t ty = 0; ty < TY; ++ty)
        for (int32_t tx = 0; tx < TX; ++tx)
          for (int32_t f = fpar; f < Nf; f += 2) {
            size_t id = ((size_t)ty * TX + tx) * Nf + f;
            if ((tx % 2) + 2 * (ty % 2) == color && lens[id] > maxlen) maxlen = lens[id];
          }
      for (int32_t step0 = 0; step0 < maxlen; step0 += m) {
        memcpy(esnap, e, sizeof(double) * ns);
#pragma omp parallel
        {
          double* ew = (double*)malloc(sizeof(double) * ns);
          double* dacc = (double*)malloc(sizeof(double) * ns);
#pragma omp for collapse(3) schedule(dynamic)
          for (int32_t ty = 0; ty < TY; ++ty)
            for (int32_t tx = 0; tx < TX; ++tx)
              for (int32_t f = 0; f < Nf; ++f) {
                if ((f % 2) != fpar || (tx % 2) + 2 * (ty % 2) != color) continue;
                size_t id = ((size_t)ty * TX + tx) * Nf + f;
                if (step0 >= lens[id]) continue;
                memcpy(ew, esnap, sizeof(double) * ns);
                memset(dacc, 0, sizeof(double) * ns);
                int32_t nv = frame_ptr[f + 1] - frame_ptr[f];
                int32_t* cnt = (int32_t*)calloc(nv + 1, sizeof(int32_t));
                int32_t* cidx = (int32_t*)calloc((size_t)(nv + 1) * g->W, sizeof(int32_t));
                double* cval = (double*)calloc((size_t)(nv + 1) * g->W, sizeof(double));
                double* th1 = (double*)malloc(sizeof(double) * Nz);
                double* th2 = (double*)malloc(sizeof(double) * Nz);
                for (int32_t step = step0; step < step0 + m && step < lens[id]; ++step) {
                  int32_t c = lists[id * T * T + step], yy = c / Nx, xx = c % Nx;
                  orc_line_columns(g, cosv, sinv, H, f, yy, xx, frame_ptr,
                                   frame_views, cnt, cidx, cval);
                  for (int32_t z = 0; z < Nz; ++z)
                    orc_thetas_cols(g, f, z, frame_ptr, frame_views, cnt, cidx,
                                    cval, ew, w, &th1[z], &th2[z]);
                  for (int par = 0; par < 2; ++par)
                    for (int32_t z = par; z < Nz; z += 2) {
                      size_t o = (((size_t)f * Nz + z) * Ny + yy) * Nx + xx;
                      double nb[27], tn[2];
                      uint8_t nbf[27], tf[2];
                      for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                          for (int dx = -1; dx <= 1; ++dx) {
                            int k = ((dz + 1) * 3 + dy + 1) * 3 + dx + 1;
                            int32_t z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
                            int in = z2 >= 0 && z2 < Nz && y2 >= 0 && y2 < Ny &&
                                     x2 >= 0 && x2 < Nx;
                            nb[k] = in ? x[(((size_t)f * Nz + z2) * Ny + y2) * Nx + x2] : 0;
                            nbf[k] = (uint8_t)in;
                          }
                      nb[13] = 0;
                      tf[0] = f > 0;
                      tn[0] = tf[0] ? x[o - vol] : 0;
                      tf[1] = f + 1 < Nf;
                      tn[1] = tf[1] ? x[o + vol] : 0;
                      double V = x[o];
                      if (V == 0.0 && iter > 1) {
                        int any = (tn[0] > 0.0 || tn[1] > 0.0);
                        for (int k = 0; k < 27 && !any; ++k)
                          if (nb[k] > 0.0) any = 1;
                        if (!any) continue;
                      }
                      double xn = orc_substitution(&gg, V, th1[z], th2[z], nb, nbf, tn, tf);
                      x[o] = xn;
                      for (int32_t s2 = frame_ptr[f]; s2 < frame_ptr[f + 1]; ++s2) {
                        int32_t p = frame_views[s2], si = s2 - frame_ptr[f];
                        for (int32_t q = 0; q < cnt[si]; ++q) {
                          double a = cval[si * g->W + q] * g->delta_t;
                          int32_t r = cidx[si * g->W + q];
                          for (int32_t k = 0; k < g->mult_z; ++k) {
                            size_t eo = ((size_t)p * g->N_r + r) * g->N_t + z * g->mult_z + k;
                            ew[eo] -= a * (xn - V);
                            dacc[eo] -= a * (xn - V);
                          }
                        }
                      }
                    }
                }
                for (size_t i = 0; i < ns; ++i)
                  if (dacc[i] != 0.0) {
#pragma omp atomic
                    e[i] += dacc[i];
                  }
                free(cnt);
                free(cidx);
                free(cval);
                free(th1);
                free(th2);
              }
          free(ew);
          free(dacc);
        }
      }
    }
  free(esnap);
  free(lists);
  free(lens);
}

/* ------------------------------------ asynchronous sweep model (study tool)
 * The GPU sweep's schedule restated in FP64 to study its convergence on the
 * CPU (no reference counterpart; it follows k_sweep / sweep_launches in
 * phasecon_amd/csrc/mbir_hip.hip).  Per equit 8 launches: 2x2 tile colour,
 * then frame parity.  A launch's work items are (tile, frame) pairs, each
 * placed on one of `groups` cache groups (the GPU's XCDs, each with its own
 * L2).  Every item visits its tile's masked columns in the order
 * k -> (pb + k pa) mod n, pa and pb hashed from the launch seed, the tile and
 * the frame as the kernel hashes them, and solves each column in two z-parity
 * halves.  All items of a launch advance one column per step (lock-step).
 * A column's theta is computed from the error sinogram minus the updates
 * its item cannot see yet: an update made at step j by another item becomes
 * visible at step j + lag_same + 1 when both items share a cache group, and
 * at step j + lag_cross + 1 otherwise (lag_cross < 0: when the launch ends);
 * an item always sees its own updates.  With lag_same = lag_cross = 0 the
 * items of one step still do not see each other (lock-step Jacobi); with one
 * item per frame per launch that is the sequential ICD in the GPU's order. */
typedef struct {
  int32_t tl, fl, grp;  /* tile index within the launch, frame index, group */
} orc_item;

static int32_t orc_gcd(int32_t a, int32_t b) {
  while (b) {
    int32_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

/* Work items of a launch of ntl tiles x nfl frames.  placement 0: the
 * tile-major grid (workgroup id = frame x ntl + tile, group = id mod
 * groups); 1: frames bound to groups -- each frame split into
 * r = groups / gcd(nfl, groups) parts (tiles t = part mod r), unit
 * u = part x nfl + frame placed on group u mod groups. */
static int32_t orc_place(int32_t placement, int32_t groups, int32_t ntl,
                         int32_t nfl, orc_item* out) {
  int32_t n = 0;
  if (placement == 0) {
    for (int32_t wid = 0; wid < ntl * nfl; ++wid) {
      out[n].tl = wid % ntl;
      out[n].fl = wid / ntl;
      out[n].grp = wid % groups;
      ++n;
    }
    return n;
  }
  const int32_t r = groups / orc_gcd(nfl, groups);
  for (int32_t u = 0; u < nfl * r; ++u)
    for (int32_t t = u / nfl; t < ntl; t += r) {
      out[n].tl = t;
      out[n].fl = u % nfl;
      out[n].grp = u % groups;
      ++n;
    }
  return n;
}

void orc_icd_equit_async(const orc_geom* g, const double* cosv,
                         const double* sinv, const float* H,
                         const int32_t* frame_ptr, const int32_t* frame_views,
                         const uint8_t* mask, const orc_async_opts* op,
                         double* x, double* e, const double* w,
                         double* stats) {
  orc_geom gg = *g;
  gg.alpha = op->alpha;
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z, Nf = g->N_time;
  const int32_t T = op->T, X = op->groups > 0 ? op->groups : 1;
  const int32_t TX = (Nx + T - 1) / T, TY = (Ny + T - 1) / T;
  const int32_t cap = g->W, zo = g->mult_z;
  const size_t vol = (size_t)Nz * Ny * Nx;
  const size_t ns = (size_t)g->N_p * g->N_r * g->N_t;
  const int lag_s = op->lag_same, lag_c = op->lag_cross;
  const int hidden = lag_s != 0 || lag_c != 0;
  int32_t nvmax = 0;
  for (int32_t f = 0; f < Nf; ++f)
    if (frame_ptr[f + 1] - frame_ptr[f] > nvmax) nvmax = frame_ptr[f + 1] - frame_ptr[f];
  /* tiles as build_tiles: masked columns in raster order, grouped by colour */
  int32_t* tptr = (int32_t*)calloc((size_t)TX * TY + 1, sizeof(int32_t));
  int32_t* tcols = (int32_t*)malloc(sizeof(int32_t) * ((size_t)Nx * Ny + 1));
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * ((size_t)TX * TY + 1));
  int32_t off[5], nt = 0, nc = 0;
  for (int32_t ty = 0; ty < TY; ++ty)
    for (int32_t tx = 0; tx < TX; ++tx) {
      for (int32_t yy = ty * T; yy < ty * T + T && yy < Ny; ++yy)
        for (int32_t xx = tx * T; xx < tx * T + T && xx < Nx; ++xx)
          if (mask[yy * Nx + xx]) tcols[nc++] = yy * Nx + xx;
      tptr[ty * TX + tx + 1] = nc;
    }
  for (int colour = 0; colour < 4; ++colour) {
    off[colour] = nt;
    for (int32_t ty = 0; ty < TY; ++ty)
      for (int32_t tx = 0; tx < TX; ++tx) {
        int32_t id = ty * TX + tx;
        if ((tx % 2) + 2 * (ty % 2) == colour && tptr[id + 1] > tptr[id]) order[nt++] = id;
      }
  }
  off[4] = nt;
  const int32_t max_items = nt * ((Nf + 1) / 2) * 1 + 1;
  orc_item* items = (orc_item*)malloc(sizeof(orc_item) * (size_t)max_items);
  /* per item: visiting order parameters and the column's update entries */
  int32_t* pa = (int32_t*)malloc(sizeof(int32_t) * (size_t)max_items);
  int32_t* pb = (int32_t*)malloc(sizeof(int32_t) * (size_t)max_items);
  const size_t ment = (size_t)Nz * nvmax * cap * zo;  /* entries per column */
  const int ring = (lag_s > (lag_c > 0 ? lag_c : 0) ? lag_s : (lag_c > 0 ? lag_c : 0)) + 1;
  size_t* didx = (size_t*)malloc(sizeof(size_t) * ring * (size_t)max_items * ment);
  double* dval = (double*)malloc(sizeof(double) * ring * (size_t)max_items * ment);
  int32_t* dcnt = (int32_t*)calloc((size_t)ring * max_items, sizeof(int32_t));
  double *hs = NULL, *hc = NULL, *htot = NULL, *own_all = NULL;
  if (hidden) {
    own_all = (double*)calloc((size_t)omp_get_max_threads() * ns, sizeof(double));
    hs = (double*)calloc((size_t)X * ns, sizeof(double));
    hc = (double*)calloc((size_t)X * ns, sizeof(double));
    htot = (double*)calloc(ns, sizeof(double));
  }
  double dsum = 0, xsum = 0, skipped = 0;
  for (int colour = 0; colour < 4; ++colour)
    for (int fpar = 0; fpar < 2; ++fpar) {
      const int32_t ntl = off[colour + 1] - off[colour];
      const int32_t nfl = (Nf - fpar + 1) / 2;
      if (ntl == 0 || nfl <= 0) continue;
      const int32_t ni = orc_place(op->placement, X, ntl, nfl, items);
      const uint32_t lseed = op->seed * 2654435761U +
                             (uint32_t)(op->call_index * 8 + colour * 2 + fpar) * 40503U;
      int32_t steps = 0;
      for (int32_t i = 0; i < ni; ++i) {
        const int32_t tile = order[off[colour] + items[i].tl];
        const int32_t f = fpar + 2 * items[i].fl;
        const int32_t n = tptr[tile + 1] - tptr[tile];
        const uint32_t h = orc_hash(lseed ^ orc_hash((uint32_t)tile * 131071U + (uint32_t)f * 524287U));
        pa[i] = 1;
        pb[i] = 0;
        if (n > 1) {
          pa[i] = 1 + (int32_t)(h % (uint32_t)(n - 1));
          while (orc_gcd(pa[i], n) != 1) pa[i] = pa[i] % (n - 1) + 1;
          pb[i] = (int32_t)(orc_hash(h + 0x9e3779b9U) % (uint32_t)n);
        }
        if (n > steps) steps = n;
      }
      memset(dcnt, 0, sizeof(int32_t) * (size_t)ring * max_items);
      for (int32_t k = 0; k < steps; ++k) {
        const int slot = k % ring;
#pragma omp parallel reduction(+ : dsum, xsum, skipped)
        {
          const int tid = omp_get_thread_num();
          double* own = hidden ? own_all + (size_t)tid * ns : NULL;
          int32_t* cnt = (int32_t*)calloc(nvmax + 1, sizeof(int32_t));
          int32_t* cidx = (int32_t*)calloc((size_t)(nvmax + 1) * cap, sizeof(int32_t));
          double* cval = (double*)calloc((size_t)(nvmax + 1) * cap, sizeof(double));
          double* th1 = (double*)malloc(sizeof(double) * Nz);
          double* th2 = (double*)malloc(sizeof(double) * Nz);
#pragma omp for schedule(dynamic, 4)
          for (int32_t i = 0; i < ni; ++i) {
            const int32_t tile = order[off[colour] + items[i].tl];
            const int32_t f = fpar + 2 * items[i].fl, grp = items[i].grp;
            const int32_t n = tptr[tile + 1] - tptr[tile];
            size_t* my_idx = didx + ((size_t)slot * max_items + i) * ment;
            double* my_val = dval + ((size_t)slot * max_items + i) * ment;
            dcnt[(size_t)slot * max_items + i] = 0;
            if (k >= n) continue;
            const int32_t c = tcols[tptr[tile] + (int32_t)(((int64_t)pb[i] + (int64_t)k * pa[i]) % n)];
            const int32_t yy = c / Nx, xx = c % Nx;
            orc_line_columns(g, cosv, sinv, H, f, yy, xx, frame_ptr, frame_views,
                             cnt, cidx, cval);
            /* own recent updates, visible to their item */
            if (hidden)
              for (int back = 1; back <= lag_s && back <= k; ++back) {
                const size_t b = ((size_t)((k - back) % ring) * max_items + i);
                for (int32_t q = 0; q < dcnt[b]; ++q) own[didx[b * ment + q]] += dval[b * ment + q];
              }
            /* theta per z from the sinogram this item sees */
            const int32_t s0 = frame_ptr[f], nv = frame_ptr[f + 1] - s0;
            for (int32_t z = 0; z < Nz; ++z) {
              double t1 = 0, t2 = 0;
              for (int32_t si = 0; si < nv; ++si) {
                const int32_t p = frame_views[s0 + si];
                for (int32_t m = 0; m < cnt[si]; ++m) {
                  const double a = cval[si * cap + m] * g->delta_t;
                  const int32_t r = cidx[si * cap + m];
                  for (int32_t kk = 0; kk < zo; ++kk) {
                    const size_t o = ((size_t)p * g->N_r + r) * g->N_t + z * zo + kk;
                    double evv = e[o];
                    if (hidden)
                      evv -= (hs[(size_t)grp * ns + o] - own[o]) + (htot[o] - hc[(size_t)grp * ns + o]);
                    const double wv = w[o];
                    if (fabs(evv * sqrt(wv)) < g->huber_T) {
                      t2 += a * a * wv;
                      t1 -= evv * a * wv;
                    } else {
                      const double hw = g->huber_T * g->huber_delta * sqrt(wv) / fabs(evv);
                      t2 += a * a * hw;
                      t1 -= evv * a * hw;
                    }
                  }
                }
              }
              th1[z] = t1;
              th2[z] = t2;
            }
            if (hidden)
              for (int back = 1; back <= lag_s && back <= k; ++back) {
                const size_t b = ((size_t)((k - back) % ring) * max_items + i);
                for (int32_t q = 0; q < dcnt[b]; ++q) own[didx[b * ment + q]] = 0.0;
              }
            /* the column's two z-parity halves */
            int32_t ne = 0;
            for (int par = 0; par < 2; ++par)
              for (int32_t z = par; z < Nz; z += 2) {
                const size_t o = (((size_t)f * Nz + z) * Ny + yy) * Nx + xx;
                double nb[27], tn[2];
                uint8_t nbf[27], tf[2];
                for (int dz = -1; dz <= 1; ++dz)
                  for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                      const int kq = ((dz + 1) * 3 + dy + 1) * 3 + dx + 1;
                      const int32_t z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
                      const int in = z2 >= 0 && z2 < Nz && y2 >= 0 && y2 < Ny && x2 >= 0 && x2 < Nx;
                      nb[kq] = in ? x[(((size_t)f * Nz + z2) * Ny + y2) * Nx + x2] : 0;
                      nbf[kq] = (uint8_t)in;
                    }
                nb[13] = 0;
                tf[0] = f > 0;
                tn[0] = tf[0] ? x[o - vol] : 0;
                tf[1] = f + 1 < Nf;
                tn[1] = tf[1] ? x[o + vol] : 0;
                const double V = x[o];
                if (V == 0.0 && op->iter > 1) {
                  int any = (tn[0] > 0.0 || tn[1] > 0.0);
                  for (int kq = 0; kq < 27 && !any; ++kq)
                    if (nb[kq] > 0.0) any = 1;
                  if (!any) {
                    skipped += 1;
                    continue;
                  }
                }
                const double xn = orc_substitution(&gg, V, th1[z], th2[z], nb, nbf, tn, tf);
                x[o] = xn;
                dsum += fabs(xn - V);
                xsum += fabs(xn);
                if (xn == V) continue;
                for (int32_t si = 0; si < nv; ++si) {
                  const int32_t p = frame_views[s0 + si];
                  for (int32_t m = 0; m < cnt[si]; ++m) {
                    const double a = cval[si * cap + m] * g->delta_t;
                    const int32_t r = cidx[si * cap + m];
                    for (int32_t kk = 0; kk < zo; ++kk) {
                      my_idx[ne] = ((size_t)p * g->N_r + r) * g->N_t + z * zo + kk;
                      my_val[ne] = -a * (xn - V);
                      ++ne;
                    }
                  }
                }
              }
            dcnt[(size_t)slot * max_items + i] = ne;
          }
          free(cnt);
          free(cidx);
          free(cval);
          free(th1);
          free(th2);
        }
        /* apply the step's updates; items of one frame touch only that
         * frame's views, so frames are applied in parallel */
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t fl = 0; fl < nfl; ++fl)
          for (int32_t i = 0; i < ni; ++i) {
            if (items[i].fl != fl) continue;
            const int32_t grp = items[i].grp;
            const size_t b = (size_t)slot * max_items + i;
            for (int32_t q = 0; q < dcnt[b]; ++q) {
              const size_t o = didx[b * ment + q];
              const double v = dval[b * ment + q];
              e[o] += v;
              if (hidden) {
                if (lag_s > 0) hs[(size_t)grp * ns + o] += v;
                if (lag_c != 0) {
                  hc[(size_t)grp * ns + o] += v;
                  htot[o] += v;
                }
              }
            }
            /* updates leaving the hidden windows */
            if (hidden && lag_s > 0 && k - lag_s >= 0) {
              const size_t bo = (size_t)((k - lag_s) % ring) * max_items + i;
              for (int32_t q = 0; q < dcnt[bo]; ++q)
                hs[(size_t)grp * ns + didx[bo * ment + q]] -= dval[bo * ment + q];
            }
            if (hidden && lag_c > 0 && k - lag_c >= 0) {
              const size_t bo = (size_t)((k - lag_c) % ring) * max_items + i;
              for (int32_t q = 0; q < dcnt[bo]; ++q) {
                hc[(size_t)grp * ns + didx[bo * ment + q]] -= dval[bo * ment + q];
                htot[didx[bo * ment + q]] -= dval[bo * ment + q];
              }
            }
          }
      }
      /* a launch ends with every update visible */
      if (hidden) {
        memset(hs, 0, sizeof(double) * (size_t)X * ns);
        memset(hc, 0, sizeof(double) * (size_t)X * ns);
        memset(htot, 0, sizeof(double) * ns);
      }
    }
  free(hs);
  free(hc);
  free(htot);
  free(own_all);
  free(didx);
  free(dval);
  free(dcnt);
  free(pa);
  free(pb);
  free(items);
  free(order);
  free(tcols);
  free(tptr);
  if (stats) {
    stats[0] = dsum;
    stats[1] = xsum;
    stats[2] = skipped;
  }
}

/* ------------------------------ PCMBIR tomographic ICD (row a14), FP64
 * The tomographic part of PCMBIR_3D, compiled out of the shipped reference
 * (ENABLE_TOMO_RECONS is never defined): two objects, delta (PhaseObject)
 * and beta (MagObject), on the MBIR A rule times the wavenumber
 * (PCMBIR_3D/XT_AMatrix.c:103), each with its own error sinogram, coupled
 * through the decorrelating transform of the prior. */

/* b(d) = rho'(d) / d of one qGGMRF potential, rho''(0) at d = 0
 * (PCMBIR_3D/XT_Prior.c:27-118). */
static double orc_pc_b(double d, double c, double sq, double sqp, double p, double q) {
  if (d == 0) return q / (sq * c);
  double ratio = pow(fabs(d), q - p) / sqp;
  double den = c + ratio;
  double mag = pow(fabs(d), q - 1);
  double deriv = (mag / (den * sq)) * (q - ((q - p) * ratio) / den);
  return (d < 0 ? -deriv : deriv) / d;
}

static double orc_pc_rho(double d, double c, double sq, double sqp, double p, double q) {
  double ad = fabs(d);
  return (pow(ad, q) / sq) / (c + pow(ad, q - p) / sqp);
}

/* Accumulates one neighbour's terms of the coupled surrogate
 * (XT_Prior.c:148-176): with D = (Vd - nd, Vb - nb), the two decorrelated
 * differences T0 D (phase prior) and T1 D (magnitude prior) give the
 * weights b0, b1; Q = T^T diag(b0, b1) T; u += w Q (nd, nb), M += w Q. */
static void orc_pc_pair(const orc_pc_prior* P, int temporal, double w, double Vd, double Vb,
                        double nd, double nb, double u[2], double M[4]) {
  const double* T = P->T;
  double Dd = Vd - nd, Db = Vb - nb;
  double d0 = T[0] * Dd + T[1] * Db, d1 = T[2] * Dd + T[3] * Db;
  double b0 = temporal ? orc_pc_b(d0, P->c_t[0], P->sq_t[0], P->sqp_t[0], P->p, P->q)
                       : orc_pc_b(d0, P->c_s[0], P->sq_s[0], P->sqp_s[0], P->p, P->q);
  double b1 = temporal ? orc_pc_b(d1, P->c_t[1], P->sq_t[1], P->sqp_t[1], P->p, P->q)
                       : orc_pc_b(d1, P->c_s[1], P->sq_s[1], P->sqp_s[1], P->p, P->q);
  /* Q[i][j] = T[0][i] b0 T[0][j] + T[1][i] b1 T[1][j] */
  double Q[4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) Q[i * 2 + j] = T[i] * b0 * T[j] + T[2 + i] * b1 * T[2 + j];
  u[0] += w * (Q[0] * nd + Q[1] * nb);
  u[1] += w * (Q[2] * nd + Q[3] * nb);
  for (int k = 0; k < 4; ++k) M[k] += w * Q[k];
}

void orc_pc_substitution(const orc_pc_prior* P, double* Vd, double* Vb, double th1d,
                         double th1b, double th2d, double th2b, const double* nbd,
                         const double* nbb, const uint8_t* nbf, const double* tnd,
                         const double* tnb, const uint8_t* tf) {
  /* FunctionalSubstitution (XT_Prior.c:137-239): the 2x2 system
   * (diag(theta2) + M) v = u + theta2 V - theta1, then over-relaxation */
  double u[2] = {0, 0}, M[4] = {0, 0, 0, 0};
  const double V0 = *Vd, V1 = *Vb;
  for (int k = 0; k < 27; ++k)
    if (k != 13 && nbf[k]) orc_pc_pair(P, 0, P->filt[k], V0, V1, nbd[k], nbb[k], u, M);
  for (int k = 0; k < 2; ++k)
    if (tf[k]) orc_pc_pair(P, 1, P->filt_t, V0, V1, tnd[k], tnb[k], u, M);
  double a = th2d + M[0], b = M[1], c = M[2], d = th2b + M[3];
  double inv = 1.0 / (a * d - c * b);
  double r0 = u[0] + th2d * V0 - th1d, r1 = u[1] + th2b * V1 - th1b;
  double nd = inv * (d * r0 - b * r1), nb = inv * (-c * r0 + a * r1);
  *Vd = V0 + P->alpha * (nd - V0);
  *Vb = V1 + P->alpha * (nb - V1);
}

double orc_pc_cost(const orc_geom* g, const orc_pc_prior* P, const double* ed,
                   const double* eb, const double* xd, const double* xb, double* out) {
  /* mu/2 sum(e_d^2 + e_b^2), the objective the ICD minimises (the
   * reference's computeCost squares mu e, XT_ICD_update.c:92-102, which
   * agrees at mu = 1), plus the decorrelated pair potentials of
   * XT_ICD_update.c:103-270 over the 13 forward spatial and the (t, t+1)
   * pairs */
  const size_t ns = (size_t)g->N_p * g->N_r * g->N_t;
  double data = 0;
  for (size_t i = 0; i < ns; ++i) data += ed[i] * ed[i] + eb[i] * eb[i];
  data *= P->mu / 2;
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z, Nf = g->N_time;
  double prior = 0;
  for (int32_t f = 0; f < Nf; ++f)
    for (int32_t z = 0; z < Nz; ++z)
      for (int32_t yy = 0; yy < Ny; ++yy)
        for (int32_t xx = 0; xx < Nx; ++xx) {
          size_t o = (((size_t)f * Nz + z) * Ny + yy) * Nx + xx;
          for (int k = 14; k < 27; ++k) {
            int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
            int32_t z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
            if (z2 < 0 || z2 >= Nz || y2 < 0 || y2 >= Ny || x2 < 0 || x2 >= Nx) continue;
            size_t o2 = (((size_t)f * Nz + z2) * Ny + y2) * Nx + x2;
            double Dd = xd[o] - xd[o2], Db = xb[o] - xb[o2];
            prior += P->filt[k] *
                     (orc_pc_rho(P->T[0] * Dd + P->T[1] * Db, P->c_s[0], P->sq_s[0], P->sqp_s[0], P->p, P->q) +
                      orc_pc_rho(P->T[2] * Dd + P->T[3] * Db, P->c_s[1], P->sq_s[1], P->sqp_s[1], P->p, P->q));
          }
          if (f + 1 < Nf && P->filt_t != 0) {
            size_t o2 = o + (size_t)Nz * Ny * Nx;
            double Dd = xd[o] - xd[o2], Db = xb[o] - xb[o2];
            prior += P->filt_t *
                     (orc_pc_rho(P->T[0] * Dd + P->T[1] * Db, P->c_t[0], P->sq_t[0], P->sqp_t[0], P->p, P->q) +
                      orc_pc_rho(P->T[2] * Dd + P->T[3] * Db, P->c_t[1], P->sq_t[1], P->sqp_t[1], P->p, P->q));
          }
        }
  if (out) {
    out[0] = data;
    out[1] = prior;
  }
  return data + prior;
}

void orc_pc_equit(const orc_geom* g, const double* cosv, const double* sinv, const float* H,
                  const int32_t* frame_ptr, const int32_t* frame_views, const uint8_t* mask,
                  const orc_pc_prior* P, int32_t T, uint32_t seed, int32_t iter,
                  int32_t call_index, double* xd, double* xb, double* ed, double* eb,
                  double* stats) {
  /* One equit of the GPU sweep's schedule (tiles, colours, frame parity,
   * hashed column order, two z-parity halves per column), run serially:
   * the voxel update of PCMBIR_3D/XT_VoxUpdate.c:23-80 (theta of both error
   * sinograms over the column, theta2 shared), the neighbourhood and zero
   * skipping of XT_VoxUpdate.c:84-230 and the coupled substitution.  Each
   * launch's items touch disjoint views and no common neighbours, so this
   * order is the deterministic GPU sweep's when every launch holds one item
   * per frame. */
  const int32_t Nx = g->N_x, Ny = g->N_y, Nz = g->N_z, Nf = g->N_time;
  const int32_t TX = (Nx + T - 1) / T, TY = (Ny + T - 1) / T, cap = g->W, zo = g->mult_z;
  const size_t vol = (size_t)Nz * Ny * Nx;
  const double kw = P->wavenumber, mu = P->mu;
  int32_t nvmax = 0;
  for (int32_t f = 0; f < Nf; ++f)
    if (frame_ptr[f + 1] - frame_ptr[f] > nvmax) nvmax = frame_ptr[f + 1] - frame_ptr[f];
  int32_t* tptr = (int32_t*)calloc((size_t)TX * TY + 1, sizeof(int32_t));
  int32_t* tcols = (int32_t*)malloc(sizeof(int32_t) * ((size_t)Nx * Ny + 1));
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * ((size_t)TX * TY + 1));
  orc_item* items = (orc_item*)malloc(sizeof(orc_item) * ((size_t)TX * TY * Nf + 1));
  int32_t* cnt = (int32_t*)calloc(nvmax + 1, sizeof(int32_t));
  int32_t* cidx = (int32_t*)calloc((size_t)(nvmax + 1) * cap, sizeof(int32_t));
  double* cval = (double*)calloc((size_t)(nvmax + 1) * cap, sizeof(double));
  double* t1d = (double*)malloc(sizeof(double) * Nz);
  double* t1b = (double*)malloc(sizeof(double) * Nz);
  int32_t off[5], nt = 0, nc = 0;
  double dsum = 0, skipped = 0;
  for (int32_t ty = 0; ty < TY; ++ty)
    for (int32_t tx = 0; tx < TX; ++tx) {
      for (int32_t yy = ty * T; yy < ty * T + T && yy < Ny; ++yy)
        for (int32_t xx = tx * T; xx < tx * T + T && xx < Nx; ++xx)
          if (mask[yy * Nx + xx]) tcols[nc++] = yy * Nx + xx;
      tptr[ty * TX + tx + 1] = nc;
    }
  for (int colour = 0; colour < 4; ++colour) {
    off[colour] = nt;
    for (int32_t ty = 0; ty < TY; ++ty)
      for (int32_t tx = 0; tx < TX; ++tx) {
        int32_t id = ty * TX + tx;
        if ((tx % 2) + 2 * (ty % 2) == colour && tptr[id + 1] > tptr[id]) order[nt++] = id;
      }
  }
  off[4] = nt;
  for (int colour = 0; colour < 4; ++colour)
    for (int fpar = 0; fpar < 2; ++fpar) {
      const int32_t ntl = off[colour + 1] - off[colour], nfl = (Nf - fpar + 1) / 2;
      if (ntl == 0 || nfl <= 0) continue;
      const int32_t ni = orc_place(0, 8, ntl, nfl, items);
      const uint32_t lseed = seed * 2654435761U +
                             (uint32_t)(call_index * 8 + colour * 2 + fpar) * 40503U;
      for (int32_t i = 0; i < ni; ++i) {
        const int32_t tile = order[off[colour] + items[i].tl];
        const int32_t f = fpar + 2 * items[i].fl;
        const int32_t n = tptr[tile + 1] - tptr[tile];
        const uint32_t h = orc_hash(lseed ^ orc_hash((uint32_t)tile * 131071U + (uint32_t)f * 524287U));
        int32_t pa = 1, pb = 0;
        if (n > 1) {
          pa = 1 + (int32_t)(h % (uint32_t)(n - 1));
          while (orc_gcd(pa, n) != 1) pa = pa % (n - 1) + 1;
          pb = (int32_t)(orc_hash(h + 0x9e3779b9U) % (uint32_t)n);
        }
        const int32_t s0 = frame_ptr[f], nv = frame_ptr[f + 1] - s0;
        for (int32_t k = 0; k < n; ++k) {
          const int32_t c = tcols[tptr[tile] + (int32_t)(((int64_t)pb + (int64_t)k * pa) % n)];
          const int32_t yy = c / Nx, xx = c % Nx;
          orc_line_columns(g, cosv, sinv, H, f, yy, xx, frame_ptr, frame_views, cnt, cidx, cval);
          double th2 = 0;
          for (int32_t z = 0; z < Nz; ++z) t1d[z] = t1b[z] = 0;
          for (int32_t si = 0; si < nv; ++si) {
            const int32_t pv = frame_views[s0 + si];
            for (int32_t m = 0; m < cnt[si]; ++m) {
              const double a = cval[si * cap + m] * kw;
              const int32_t r = cidx[si * cap + m];
              th2 += mu * a * a * zo;
              for (int32_t z = 0; z < Nz; ++z)
                for (int32_t kk = 0; kk < zo; ++kk) {
                  const size_t o = ((size_t)pv * g->N_r + r) * g->N_t + z * zo + kk;
                  t1d[z] -= ed[o] * a * mu;
                  t1b[z] -= eb[o] * a * mu;
                }
            }
          }
          for (int par = 0; par < 2; ++par)
            for (int32_t z = par; z < Nz; z += 2) {
              const size_t o = (((size_t)f * Nz + z) * Ny + yy) * Nx + xx;
              double nbd[27], nbb[27], tnd[2], tnb[2];
              uint8_t nbf[27], tf[2];
              int any = 0;
              for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                  for (int dx = -1; dx <= 1; ++dx) {
                    const int kq = ((dz + 1) * 3 + dy + 1) * 3 + dx + 1;
                    const int32_t z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
                    const int in = z2 >= 0 && z2 < Nz && y2 >= 0 && y2 < Ny && x2 >= 0 && x2 < Nx;
                    const size_t o2 = (((size_t)f * Nz + z2) * Ny + y2) * Nx + x2;
                    nbd[kq] = in ? xd[o2] : 0;
                    nbb[kq] = in ? xb[o2] : 0;
                    nbf[kq] = (uint8_t)in;
                  }
              nbd[13] = nbb[13] = 0;
              tf[0] = f > 0;
              tf[1] = f + 1 < Nf;
              tnd[0] = tf[0] ? xd[o - vol] : 0;
              tnb[0] = tf[0] ? xb[o - vol] : 0;
              tnd[1] = tf[1] ? xd[o + vol] : 0;
              tnb[1] = tf[1] ? xb[o + vol] : 0;
              const double Vd = xd[o], Vb = xb[o];
              /* zero skipping (XT_VoxUpdate.c:208-228) */
              if (Vd == 0.0 && Vb == 0.0 && iter > 1) {
                any = tnd[0] > 0 || tnd[1] > 0 || tnb[0] > 0 || tnb[1] > 0;
                for (int kq = 0; kq < 27 && !any; ++kq)
                  if (nbd[kq] > 0 || nbb[kq] > 0) any = 1;
                if (!any) {
                  skipped += 1;
                  continue;
                }
              }
              double nd = Vd, nb = Vb;
              orc_pc_substitution(P, &nd, &nb, t1d[z], t1b[z], th2, th2, nbd, nbb, nbf, tnd, tnb, tf);
              xd[o] = nd;
              xb[o] = nb;
              dsum += sqrt((nd - Vd) * (nd - Vd) + (nb - Vb) * (nb - Vb));
              for (int32_t si = 0; si < nv; ++si) {
                const int32_t pv = frame_views[s0 + si];
                for (int32_t m = 0; m < cnt[si]; ++m) {
                  const double a = cval[si * cap + m] * kw;
                  const int32_t r = cidx[si * cap + m];
                  for (int32_t kk = 0; kk < zo; ++kk) {
                    const size_t oe = ((size_t)pv * g->N_r + r) * g->N_t + z * zo + kk;
                    ed[oe] -= a * (nd - Vd);
                    eb[oe] -= a * (nb - Vb);
                  }
                }
              }
            }
        }
      }
    }
  free(tptr);
  free(tcols);
  free(order);
  free(items);
  free(cnt);
  free(cidx);
  free(cval);
  free(t1d);
  free(t1b);
  if (stats) {
    stats[0] = dsum;
    stats[1] = 0;
    stats[2] = skipped;
  }
}

void orc_sino_from_file(int32_t N_p, int32_t N_t, int32_t N_r,
                        const float* in, double* out) {
  /* XT_Init.c:294-301: file index p*N_t*N_r + t*N_r + r */
  for (int32_t p = 0; p < N_p; ++p)
    for (int32_t r = 0; r < N_r; ++r)
      for (int32_t t = 0; t < N_t; ++t)
        out[((size_t)p * N_r + r) * N_t + t] = in[((size_t)p * N_t + t) * N_r + r];
}


/* ------------------------------------------------------------------ ring
 * offset and noise variance */

int32_t orc_offset_windows(int32_t N_r, int32_t N_t, double* win) {
  /* window sizes and steps of XT_OffsetError.c:57-61 (integer arithmetic) */
  const int32_t rs = 2 * N_r / (int32_t)(sqrt((double)N_r) + 0.5);
  const int32_t ts = 2 * N_t / (int32_t)(sqrt((double)N_t) + 0.5);
  int32_t m = 0;
  for (int32_t i = 0; i <= N_r - rs / 2; i += rs / 2)
    for (int32_t j = 0; j <= N_t - ts / 2; j += ts / 2) {
      if (win) {
        double* wm = win + (size_t)m * N_r * N_t;
        /* triangular weights, wrapped around both axes; a wrapped sample
         * overwrites (XT_OffsetError.c:68-79) */
        for (int32_t k = i; k < i + rs; ++k)
          for (int32_t l = j; l < j + ts; ++l) {
            const double a = (k - i) < rs / 2 ? (k - i + 1) : rs - (k - i);
            const double b = (l - j) < ts / 2 ? (l - j + 1) : ts - (l - j);
            wm[(size_t)(k % N_r) * N_t + (l % N_t)] = a * b;
          }
      }
      ++m;
    }
  return m;
}

/* inverse of an n x n matrix by Gauss-Jordan elimination with partial
 * pivoting (in place, a is n x n row-major); returns -1 if singular */
static int orc_invert(double* a, int32_t n) {
  double* inv = (double*)calloc((size_t)n * n, sizeof(double));
  for (int32_t i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int32_t c = 0; c < n; ++c) {
    int32_t piv = c;
    for (int32_t r = c + 1; r < n; ++r)
      if (fabs(a[(size_t)r * n + c]) > fabs(a[(size_t)piv * n + c])) piv = r;
    if (a[(size_t)piv * n + c] == 0.0) {
      free(inv);
      return -1;
    }
    if (piv != c)
      for (int32_t k = 0; k < n; ++k) {
        double t = a[(size_t)c * n + k];
        a[(size_t)c * n + k] = a[(size_t)piv * n + k];
        a[(size_t)piv * n + k] = t;
        t = inv[(size_t)c * n + k];
        inv[(size_t)c * n + k] = inv[(size_t)piv * n + k];
        inv[(size_t)piv * n + k] = t;
      }
    const double p = a[(size_t)c * n + c];
    for (int32_t k = 0; k < n; ++k) {
      a[(size_t)c * n + k] /= p;
      inv[(size_t)c * n + k] /= p;
    }
    for (int32_t r = 0; r < n; ++r)
      if (r != c) {
        const double f = a[(size_t)r * n + c];
        if (f == 0.0) continue;
        for (int32_t k = 0; k < n; ++k) {
          a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
          inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
        }
      }
  }
  memcpy(a, inv, sizeof(double) * (size_t)n * n);
  free(inv);
  return 0;
}

/* per (r, t): the generalized-Huber weight of every view's entry -- w where
 * selected, T delta sqrt(w) / |e| otherwise -- summed (lam), and the matching
 * numerator of the offset step (the mode-3 right-hand side when rhs3) */
static void orc_offset_sums(int32_t N_p, int32_t N_r, int32_t N_t, double T, double delta,
                            const double* w, const double* e, const double* d,
                            const uint8_t* sel, int rhs3, double* num, double* lam) {
  const size_t nrt = (size_t)N_r * N_t;
  for (size_t i = 0; i < nrt; ++i) {
    double a = 0.0, b = 0.0;
    for (int32_t p = 0; p < N_p; ++p) {
      const size_t q = (size_t)p * nrt + i;
      const double tw = T * delta * sqrt(w[q]);
      if (sel[q]) {
        b += w[q];
        a += (rhs3 ? e[q] + d[i] : e[q]) * w[q];
      } else {
        const double l = tw / fabs(e[q]);
        b += l;
        a += rhs3 ? (e[q] + d[i]) * l : tw * (double)((e[q] > 0) - (e[q] < 0));
      }
    }
    num[i] = a;
    lam[i] = b;
  }
}

static void orc_reselect(size_t n, double T, const double* w, const double* e, uint8_t* sel) {
  for (size_t q = 0; q < n; ++q) sel[q] = fabs(e[q] * sqrt(w[q])) < T;
}

void orc_offset_update(int32_t mode, int32_t N_p, int32_t N_r, int32_t N_t,
                       double T, double delta, const double* w, double* e,
                       double* d, uint8_t* sel) {
  const size_t nrt = (size_t)N_r * N_t, n = nrt * N_p;
  double* num = (double*)malloc(sizeof(double) * nrt);
  double* lam = (double*)malloc(sizeof(double) * nrt);
  double* step = (double*)malloc(sizeof(double) * nrt);
  orc_offset_sums(N_p, N_r, N_t, T, delta, w, e, d, sel, mode == 3, num, lam);
  if (mode == 1) {
    for (size_t i = 0; i < nrt; ++i) step[i] = num[i] / lam[i];
  } else if (mode == 2) {
    /* the Lagrange multiplier of sum(d) = 0 */
    double ns = 0.0, ds = 0.0;
    for (size_t i = 0; i < nrt; ++i) {
      ns += d[i] + num[i] / lam[i];
      ds += 1.0 / lam[i];
    }
    const double gamma = ns / ds;
    for (size_t i = 0; i < nrt; ++i) step[i] = (num[i] - gamma) / lam[i];
  } else {
    /* minimise sum lam (x - b/lam)^2 subject to <A_m, x> = 0 for every
     * window m: x = (b - sum_m A_m c_m) / lam, c = D^-1 (A b / lam),
     * D = A diag(1/lam) A^T */
    const int32_t M = orc_offset_windows(N_r, N_t, NULL);
    double* A = (double*)calloc((size_t)M * nrt, sizeof(double));
    double* D = (double*)calloc((size_t)M * M, sizeof(double));
    double* v = (double*)calloc(M, sizeof(double));
    double* c = (double*)calloc(M, sizeof(double));
    orc_offset_windows(N_r, N_t, A);
    for (int32_t i = 0; i < M; ++i)
      for (int32_t j = 0; j < M; ++j) {
        double s = 0.0;
        for (size_t k = 0; k < nrt; ++k) s += A[(size_t)i * nrt + k] * A[(size_t)j * nrt + k] / lam[k];
        D[(size_t)i * M + j] = s;
      }
    orc_invert(D, M);
    for (int32_t i = 0; i < M; ++i) {
      double s = 0.0;
      for (size_t k = 0; k < nrt; ++k) s += A[(size_t)i * nrt + k] * num[k] / lam[k];
      v[i] = s;
    }
    for (int32_t i = 0; i < M; ++i) {
      double s = 0.0;
      for (int32_t j = 0; j < M; ++j) s += D[(size_t)i * M + j] * v[j];
      c[i] = s;
    }
    for (size_t k = 0; k < nrt; ++k) {
      double s = 0.0;
      for (int32_t m = 0; m < M; ++m) s += A[(size_t)m * nrt + k] * c[m];
      /* new offset x; the step applied to e and d is x - d */
      step[k] = (num[k] - s) / lam[k] - d[k];
    }
    free(A);
    free(D);
    free(v);
    free(c);
  }
  for (size_t i = 0; i < nrt; ++i) d[i] += step[i];
  for (size_t q = 0; q < n; ++q) e[q] -= step[q % nrt];
  orc_reselect(n, T, w, e, sel);
  free(num);
  free(lam);
  free(step);
}

void orc_variance_update(int64_t n, double T, double delta, double* w,
                         const double* e, uint8_t* sel, double* var) {
  double s = 0.0;
  for (int64_t q = 0; q < n; ++q) {
    w[q] *= *var; /* back to the measured weights */
    s += sel[q] ? e[q] * e[q] * w[q] : fabs(e[q]) * delta * T * sqrt(w[q] * *var);
  }
  *var = s / (double)n;
  for (int64_t q = 0; q < n; ++q) w[q] /= *var;
  orc_reselect((size_t)n, T, w, e, sel);
}

