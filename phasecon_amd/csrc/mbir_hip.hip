// mbir_hip.hip -- gfx950 kernels of the MBIR hot path and the thin C-ABI
// declared in include/mbir_hip.h.
//
// Device data layout (HBM-resident for the whole stage):
//   y, w, e       [N_p][N_r][N_t] float      detector row t fastest, so one
//                                            voxel column's rows are contiguous
//   x             [N_time][N_y][N_x][N_z]    z fastest: a voxel column and its
//                                            3x3 neighbour columns are 9
//                                            contiguous runs of N_z floats
//   A             first[col][slot] int32, val[col][slot][W] float, slots in
//                 frame order, so one (column, frame) visit reads one
//                 contiguous run of P_f*W values.
// Kernels:
//   k_amatrix     A build, FP64 arithmetic identical to the reference rule
//   k_fwdproj     A x per view with an LDS accumulator (FP32 for e, FP64 for
//                 the cost), A recomputed on the fly in FP64
//   k_sweep       the parallel ICD: one wavefront per (tile, frame) walks the
//                 tile's voxel columns in a seeded random order; theta1/theta2
//                 reduced over the column's sparse footprint across the
//                 wavefront, qGGMRF surrogate from an LDS-staged 3x3xN_z (+t+-1)
//                 neighbourhood, error sinogram updated with atomics
//   k_prior_cost  FP64 prior term, k_data_cost FP64 data term
//   k_backproj    A^T v, k_thetas / k_subst unit-test entry kernels
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <algorithm>
#include <vector>
#include <mutex>

#include "mbir_hip.h"
#include "mbir_engine.h"

namespace {

constexpr int kNB = 65;        // detector response entries per view
constexpr int kRespBins = 64;  // XT_Constants.h:79
constexpr int kNZMax = 128;    // voxels per column (N_z) supported by k_sweep
constexpr int kNZPc = 256;     // voxels per column supported by the tomographic sweep
constexpr int kNTMax = 256;    // detector rows (N_t) supported by k_sweep
constexpr int kWMax = 33;      // window width 2*mult_xy+1 (mult_xy <= 16)
constexpr int kWave = 64;
constexpr size_t kSweepXMax = 48 * 1024;  // LDS bytes for a sweep tile slab

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

#define HIP_OK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// Runs CALL with the compile-time window width kW = W (2 mult_xy + 1) of
// the sweep kernels.
#define MBIR_DISPATCH_W(W, CALL)                                                  \
  switch (W) {                                                                    \
    case 3: { constexpr int kW = 3; CALL; } break;                                \
    case 5: { constexpr int kW = 5; CALL; } break;                                \
    case 7: { constexpr int kW = 7; CALL; } break;                                \
    case 9: { constexpr int kW = 9; CALL; } break;                                \
    case 17: { constexpr int kW = 17; CALL; } break;                              \
    case 33: { constexpr int kW = 33; CALL; } break;                              \
    default: return fail("mult_xy must be 1, 2, 3, 4, 8 or 16 for the sweep");    \
  }

// Parameters passed to kernels by value.
struct DevGeom {
  int N_r, N_t, N_p, N_x, N_y, N_z, N_time, W, zo, nslots;
  double delta_r, R0, RMax, x0, y0, delta_xy, OffsetR;
  float delta_t;
  float q, p_s, p_t, c_s, c_t;
  float inv_sq_s, inv_sqp_s, inv_sq_t, inv_sqp_t;  // 1/sigma^q, 1/sigma^(q-p)
  float b0_s, b0_t;                                // rho''(0) = q/(sigma^q c)
  float filt[27];
  float filt_t;
  float alpha;
  float huber_T, huber_delta;
  double dq, dp_s, dp_t, dc_s, dc_t, dsq_s, dsqp_s, dsq_t, dsqp_t;
  double dfilt[27];
  double dfilt_t, dhuber_T, dhuber_delta;
};

__device__ __forceinline__ unsigned mix32(unsigned a) {
  a ^= a >> 16;
  a *= 0x7feb352dU;
  a ^= a >> 15;
  a *= 0x846ca68bU;
  a ^= a >> 16;
  return a;
}

__device__ __forceinline__ int gcd_i(int a, int b) {
  while (b) {
    int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// ---------------------------------------------------------------- A rule
// The A window of column (row, col) at one view: XT_AMatrix.c:50-110.  The
// arithmetic is FP64 in the reference's operation order, so indices match
// exactly and values to the last bit before any float store.  Calls
// fn(j, j - jlo, value) for each kept bin j (f1 > 0) and returns the clamped
// lower bin jlo, or -1 if the voxel misses the detector.
template <typename Fn>
__device__ __forceinline__ int amatrix_visit(const DevGeom& G, const float* __restrict__ h,
                                             double cs, double sn, int row, int col,
                                             Fn&& fn) {
  const double xc = G.x0 + ((double)col + 0.5) * G.delta_xy;
  const double yc = G.y0 + ((double)row + 0.5) * G.delta_xy;
  const double r = xc * cs - yc * sn;
  const double rlo = r - G.delta_xy, rhi = r + G.delta_xy;
  if (rhi < G.R0 || rlo > G.RMax) return -1;
  int jlo = (int)(floor(((rlo - G.R0) / G.delta_r)));
  int jhi = (int)(floor((rhi - G.R0) / G.delta_r));
  if (jhi >= G.N_r) jhi = G.N_r - 1;
  if (jlo < 0) jlo = 0;
  for (int j = jlo; j <= jhi && j - jlo < G.W; ++j) {
    const double centre = (G.R0 + (((double)j) + 0.5) * (G.delta_r));
    const double d = fabs(r - centre);
    const int k = (int)(floor((d / G.OffsetR)));
    if (k >= 0 && k < kRespBins) {
      const double w_hi = d - k * G.OffsetR;
      const double w_lo = G.OffsetR - w_hi;
      const double f = (w_lo * h[k] + w_hi * h[k + 1]) / G.OffsetR;
      if (f > 0) fn(j, j - jlo, f);
    }
  }
  return jlo;
}

__global__ __launch_bounds__(256) void k_amatrix(DevGeom G,
                                                 const float* __restrict__ H,
                                                 const double* __restrict__ cosv,
                                                 const double* __restrict__ sinv,
                                                 const int* __restrict__ slot_view,
                                                 int* __restrict__ afirst,
                                                 float* __restrict__ aval) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)G.N_x * G.N_y * G.nslots;
  if (i >= total) return;
  const int s = (int)(i % G.nslots);
  const int c = (int)(i / G.nslots);
  const int p = slot_view[s];
  float* dst = aval + i * G.W;
  for (int k = 0; k < G.W; ++k) dst[k] = 0.f;
  int kept = 0;
  const int jlo = amatrix_visit(G, H + (size_t)p * kNB, cosv[p], sinv[p], c / G.N_x, c % G.N_x,
                                [&](int, int k, double v) {
                                  dst[k] = (float)v;
                                  ++kept;
                                });
  // a column with no kept bin is empty (count 0 in XT_AMatrix.c:109)
  afirst[i] = kept ? jlo : -1;
}

// FP16 copy of A with one FP32 scale per voxel column, the largest |a| of the
// column over all view slots: q = fp16(a / s), rounded to nearest even, and
// the FP32 copy rewritten to q s, so that every kernel that reads stored A
// uses the values the FP16 sweep converts back.  One wavefront per column.
__global__ __launch_bounds__(256) void k_aquant(int ncol, int per_col, float* __restrict__ aval,
                                                unsigned short* __restrict__ aval16,
                                                float* __restrict__ scale) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncol) return;
  const int lane = threadIdx.x & 63;
  float* a = aval + (size_t)c * per_col;
  unsigned short* h = aval16 + (size_t)c * per_col;
  float m = 0.f;
  for (int j = lane; j < per_col; j += 64) m = fmaxf(m, fabsf(a[j]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (lane == 0) scale[c] = m;
  for (int j = lane; j < per_col; j += 64) {
    const _Float16 q = m > 0.f ? (_Float16)(a[j] / m) : (_Float16)0.f;
    h[j] = __builtin_bit_cast(unsigned short, q);
    a[j] = (float)q * m;
  }
}

// ----------------------------------------------------------- projection
// One workgroup per (view, row chunk): accumulate A x of the whole object
// for this view in LDS, then write e = y - Ax - d (mode 0), the FP64 data
// cost of y - Ax - d (mode 1), or Ax itself (mode 2).
template <typename Acc>
__global__ __launch_bounds__(256) void k_fwdproj(
    DevGeom G, int mode, int tchunk, const float* __restrict__ H,
    const double* __restrict__ cosv, const double* __restrict__ sinv,
    const int* __restrict__ view_frame, const float* __restrict__ x,
    const float* __restrict__ y, const float* __restrict__ w,
    const float* __restrict__ d, float* __restrict__ out,
    double* __restrict__ partial, double xscale) {
  extern __shared__ __align__(16) float smem[];
  Acc* acc = reinterpret_cast<Acc*>(smem);
  __shared__ float sh[kNB];
  __shared__ double red[256];
  const int p = blockIdx.x;
  const int t0 = blockIdx.y * tchunk;
  const int tc = min(tchunk, G.N_t - t0);
  const int tid = threadIdx.x;
  for (int i = tid; i < G.N_r * tc; i += 256) acc[i] = Acc(0);
  for (int i = tid; i < kNB; i += 256) sh[i] = H[(size_t)p * kNB + i];
  __syncthreads();
  const int f = view_frame[p];
  if (f >= 0) {
    const double cs = cosv[p], sn = sinv[p];
    const size_t vol = (size_t)G.N_y * G.N_x * G.N_z;
    const float* xf = x + (size_t)f * vol;
    const int ncol = G.N_x * G.N_y;
    for (int c = tid; c < ncol; c += 256) {
      const float* xc = xf + (size_t)c * G.N_z;
      bool any = false;
      for (int z = t0 / G.zo; z <= (t0 + tc - 1) / G.zo; ++z) any |= (xc[z] != 0.f);
      if (!any) continue;
      amatrix_visit(G, sh, cs, sn, c / G.N_x, c % G.N_x, [&](int j, int, double v) {
        Acc* row = acc + (size_t)j * tc;
        for (int t = 0; t < tc; ++t) {
          const float xv = xc[(t0 + t) / G.zo];
          if (xv != 0.f) atomicAdd(&row[t], (Acc)((double)xv * v * xscale));
        }
      });
    }
  }
  __syncthreads();
  double local = 0.0;
  for (int i = tid; i < G.N_r * tc; i += 256) {
    const int r = i / tc, t = t0 + i % tc;
    const size_t o = ((size_t)p * G.N_r + r) * G.N_t + t;
    const float dv = d ? d[(size_t)r * G.N_t + t] : 0.f;
    if (mode == 0) {
      // XT_ICD_update.c:689, rounded once to the float array
      out[o] = (float)((double)y[o] - (double)acc[i] - (double)dv);
    } else if (mode == 1) {
      const double ev = (double)y[o] - (double)acc[i] - (double)dv;
      const double s = ev * sqrt((double)w[o]);
      if (fabs(s) < G.dhuber_T)
        local += s * s;
      else
        local += 2.0 * G.dhuber_delta * G.dhuber_T * fabs(s) +
                 G.dhuber_T * G.dhuber_T * (1.0 - 2.0 * G.dhuber_delta);
    } else {
      out[o] = (float)acc[i];
    }
  }
  if (mode == 1) {
    red[tid] = local;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) partial[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = red[0];
  }
}

// Data term from the maintained error sinogram (XT_ICD_update.c:192-204).
__global__ __launch_bounds__(256) void k_data_cost(DevGeom G, size_t n,
                                                   const float* __restrict__ e,
                                                   const float* __restrict__ w,
                                                   double* __restrict__ partial) {
  __shared__ double red[256];
  double local = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (size_t)gridDim.x * 256) {
    const double s = (double)e[i] * sqrt((double)w[i]);
    if (fabs(s) < G.dhuber_T)
      local += s * s;
    else
      local += 2.0 * G.dhuber_delta * G.dhuber_T * fabs(s) +
               G.dhuber_T * G.dhuber_T * (1.0 - 2.0 * G.dhuber_delta);
  }
  red[threadIdx.x] = local;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// qGGMRF potential in FP64 (XT_ICD_update.c:82-90).
__device__ __forceinline__ double qg_value64(double delta, double q, double p,
                                             double c, double sq, double sqp) {
  const double ad = fabs(delta);
  if (ad == 0.0) return 0.0;
  const double lg = log2(ad);
  const double aq = (q == 2.0) ? ad * ad : exp2(q * lg);
  return (aq / sq) / (c + exp2((q - p) * lg) / sqp);
}

// qGGMRF potential with q = 2 evaluated in FP32 (v_log_f32 / v_exp_f32 /
// v_rcp_f32, a few ulp per term) for the cost estimate of kFast.
__device__ __forceinline__ float qg_value32(float delta, float qmp, float c, float inv_sq,
                                            float inv_sqp) {
  const float ad = fabsf(delta);
  const float r = __builtin_amdgcn_exp2f(qmp * __builtin_amdgcn_logf(ad)) * inv_sqp;
  const float v = ad * ad * inv_sq * __builtin_amdgcn_rcpf(c + r);
  return ad == 0.f ? 0.f : v;
}

// Prior term: each neighbour pair once (13 forward spatial + t+1)
// (XT_ICD_update.c:205-298), accumulated in FP64.  Each pair's potential is
// FP64 (kFast = false) or FP32 (kFast = true, an estimate within about 1e-6
// relative, for stopping tests that an FP64 evaluation then confirms).
template <bool kFast>
__global__ __launch_bounds__(256) void k_prior_cost(DevGeom G,
                                                    const float* __restrict__ x,
                                                    double* __restrict__ partial) {
  __shared__ double red[256];
  constexpr int fwd[13][3] = {{0, 0, 1},  {0, 1, -1},  {0, 1, 0}, {0, 1, 1},
                              {1, -1, 0}, {1, 0, 0},   {1, 1, 0}, {1, -1, -1},
                              {1, -1, 1}, {1, 0, -1},  {1, 1, -1}, {1, 1, 1},
                              {1, 0, 1}};
  // one thread per voxel (consecutive threads walk a column's z, so the
  // neighbour loads coalesce); 32-bit index arithmetic (the voxel count is
  // below 2^31, checked on the host)
  const unsigned Nz = (unsigned)G.N_z, Nx = (unsigned)G.N_x;
  const unsigned ncol = (unsigned)G.N_y * Nx;
  const unsigned vol = ncol * Nz;
  const unsigned n = vol * (unsigned)G.N_time;
  double local = 0.0;
  for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const unsigned col = i / Nz, z = i - col * Nz;
    const unsigned f = col / ncol, c = col - f * ncol;
    const int yy = (int)(c / Nx), xx = (int)(c - (unsigned)yy * Nx);
    const double v = x[i];
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      const int z2 = (int)z + fwd[k][0], y2 = yy + fwd[k][1], x2 = xx + fwd[k][2];
      if (z2 >= G.N_z || y2 < 0 || y2 >= G.N_y || x2 < 0 || x2 >= G.N_x) continue;
      const double nv = x[f * vol + ((unsigned)y2 * Nx + (unsigned)x2) * Nz + (unsigned)z2];
      const int fi = ((fwd[k][0] + 1) * 3 + (fwd[k][1] + 1)) * 3 + fwd[k][2] + 1;
      if (kFast)
        local += (double)(G.filt[fi] * qg_value32((float)(v - nv), G.q - G.p_s, G.c_s,
                                                  G.inv_sq_s, G.inv_sqp_s));
      else
        local += G.dfilt[fi] * qg_value64(v - nv, G.dq, G.dp_s, G.dc_s, G.dsq_s, G.dsqp_s);
    }
    if (f + 1 < (unsigned)G.N_time) {
      const double d = v - (double)x[i + vol];
      if (kFast)
        local += (double)(G.filt_t *
                          qg_value32((float)d, G.q - G.p_t, G.c_t, G.inv_sq_t, G.inv_sqp_t));
      else
        local += G.dfilt_t * qg_value64(d, G.dq, G.dp_t, G.dc_t, G.dsq_t, G.dsqp_t);
    }
  }
  red[threadIdx.x] = local;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// --------------------------------------------------------- ICD building blocks
// qGGMRF b(delta) = rho'(delta)/delta, FP32, with q = 2 (MRF_Q,
// XT_Constants.h; XT_ICD_update.c:92-128): with r = |d|^(2-p)/sigma^(2-p),
// b = (2 - (2-p) r/(c+r)) / ((c+r) sigma^2), and b(0) = 2/(sigma^2 c).
// |d|^(2-p) is one v_log_f32 / v_exp_f32 pair and the division a v_rcp_f32
// (each within about 1 ulp).
__device__ __forceinline__ float qg_b(float d, float qmp, float c, float inv_sq, float inv_sqp,
                                      float b0) {
  const float ad = fabsf(d);
  const float lg = __builtin_amdgcn_logf(ad);
  const float ratio = __builtin_amdgcn_exp2f(qmp * lg) * inv_sqp;
  const float rden = __builtin_amdgcn_rcpf(c + ratio);
  const float b = (2.f - qmp * ratio * rden) * inv_sq * rden;
  return ad == 0.f ? b0 : b;
}

// Prior constants of the ICD kernels: per prior (0 spatial, 1 temporal)
// 2 - p, c, 1/sigma^2, 1/sigma^(2-p) and b(0), then the 27 spatial
// neighbour weights ([dz][dy][dx]) and the temporal one.  Kernels copy it
// into LDS once; values read from there live in vector registers, which
// keeps them out of the scalar register file.
struct PriorTab {
  float qmp[2], c[2], inv_sq[2], inv_sqp[2], b0[2];
  float filt[28];
};

// Geometry of the ICD kernels (a subset of the stage geometry).
struct SwGeom {
  int N_r, N_t, N_x, N_y, N_z, N_time, zo, nslots;
  float delta_t, huber_T, huber_delta;
  PriorTab prior;
};

__device__ __forceinline__ void load_prior(const SwGeom& G, PriorTab* dst) {
  const float* src = reinterpret_cast<const float*>(&G.prior);
  float* d = reinterpret_cast<float*>(dst);
  for (int j = threadIdx.x; j < (int)(sizeof(PriorTab) / sizeof(float)); j += blockDim.x)
    d[j] = src[j];
}

// Weight of neighbour j of a voxel with value V: j < 27 spatial ([dz][dy][dx],
// 13 is the centre and never used), 27/28 = t-1/t+1.
__device__ __forceinline__ float nb_weight(const PriorTab& P, int j, float V, float nbv) {
  const int t = j >= 27 ? 1 : 0;
  return P.filt[j < 27 ? j : 27] *
         qg_b(V - nbv, P.qmp[t], P.c[t], P.inv_sq[t], P.inv_sqp[t], P.b0[t]);
}

// Surrogate solve and over-relaxation (XT_ICD_update.c:174-176).
__device__ __forceinline__ float substitute(float alpha, float V, float th1, float th2,
                                            float num, float den) {
  const float u = (num + th2 * V - th1) * __builtin_amdgcn_rcpf(den + th2);
  return V + alpha * (u - V);
}

// theta's reads of the shared error sinogram: workgroup-scope loads, which
// skip the CU's L1 and are served by the XCD's L2.  An L2 copy of a line may
// predate atomics performed since (by other XCDs' tiles, or at the memory
// side), so theta can miss recent updates of concurrent tiles -- the
// asynchronous staleness the parallel ICD already has, bounded by one launch
// (a launch starts from coherent caches).  The sinogram itself stays exact:
// every update is an agent-scope atomic.  Agent-scope loads (which refetch
// past the L2) cost 10 % more time per launch for the same convergence.
__device__ __forceinline__ float e_load(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void e_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-scope ordering: one wavefront per workgroup, and a wavefront's LDS
// operations execute in issue order, so a compiler fence at wavefront scope
// is enough between LDS phases (no s_waitcnt on in-flight global loads).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over aligned groups of n consecutive lanes (n a power of two <= 64),
// the result in every lane of the group: DPP quad permutes and row mirrors
// within 16-lane rows (VALU operand modifiers, no LDS traffic), lane
// shuffles across rows.
__device__ __forceinline__ float dpp_xor1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_half_mirror(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0x141, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_mirror(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0x140, 0xF, 0xF, false));
}
__device__ __forceinline__ float group_sum(float v, int n) {
  if (n >= 2) v += dpp_xor1(v);
  if (n >= 4) v += dpp_xor2(v);
  if (n >= 8) v += dpp_half_mirror(v);
  if (n >= 16) v += dpp_mirror(v);
  if (n >= 32) v += __shfl_xor(v, 16);
  if (n >= 64) v += __shfl_xor(v, 32);
  return v;
}
__device__ __forceinline__ int group_or(int v, int n) {
  const auto f = [](int a) { return __builtin_bit_cast(float, a); };
  const auto i = [](float a) { return __builtin_bit_cast(int, a); };
  if (n >= 2) v |= i(dpp_xor1(f(v)));
  if (n >= 4) v |= i(dpp_xor2(f(v)));
  if (n >= 8) v |= i(dpp_half_mirror(f(v)));
  if (n >= 16) v |= i(dpp_mirror(f(v)));
  if (n >= 32) v |= __shfl_xor(v, 16);
  if (n >= 64) v |= __shfl_xor(v, 32);
  return v;
}

struct ColumnArgs {
  const float* aval;
  const unsigned short* aval16;  // FP16 A (bits) and its per-column scale,
  const float* ascale;           // read by the kHalf sweep only
  const int* afirst;
  const int* frame_views;
  const float* w;
  float* e;
};

constexpr int kMaxViews = 512;  // views per frame held in LDS
constexpr int kMaxAV = 1024;    // A values per (column, frame) held in LDS
constexpr int kPreA = kMaxAV / kWave;
constexpr int kPreI = kMaxViews / kWave;
constexpr int kTnPre = 2 * kNZMax / kWave;  // t-1/t+1 column values per lane

// One (column, frame) A window in LDS: per view slot the error-sinogram
// offset of the window's first bin at row 0 (-1 where the column misses the
// detector in that view), and the W values of each slot.
struct Win {
  int* base;
  float* av;
};

// Dynamic LDS of the sweep: the tile slab, three A windows (the previous,
// current and next column), the frame's view list and the tile's column list
// (sizes in 4-byte words).
constexpr int kWinBufs = 3;
__host__ __device__ inline size_t sweep_lds_words(size_t xfloats, int nvmax, int W, int ncols) {
  return xfloats + kWinBufs * ((size_t)nvmax + (size_t)nvmax * W) + (size_t)nvmax +
         (size_t)ncols;
}

struct SweepLds {
  float* x;
  Win win[kWinBufs];
  int* views;
  int* cols;
  int wstride;  // words between consecutive window buffers
  __device__ SweepLds(float* base, int xfloats, int nvmax, int W) {
    x = base;
    wstride = nvmax * (1 + W);
    float* p = base + xfloats;
    for (int b = 0; b < kWinBufs; ++b) {
      win[b].base = reinterpret_cast<int*>(p);
      p += nvmax;
      win[b].av = p;
      p += nvmax * W;
    }
    views = reinterpret_cast<int*>(p);
    p += nvmax;
    cols = reinterpret_cast<int*>(p);
  }
};

// Register staging of a column's A window and t+-1 values (prefetch): PI
// first bins and PA values per lane (kSmall: frames of at most 64 views and
// 256 window values).  kHalf: the values stay FP16 bits until they are
// stored, so that no conversion waits on the loads in flight.
template <bool kSmall, bool kHalf = false>
struct APre {
  static constexpr int PI = kSmall ? 1 : kPreI;
  static constexpr int PA = kSmall ? 4 : kPreA;
  using AV = typename std::conditional<kHalf, unsigned short, float>::type;
  int a0[PI];
  AV av[PA];
  float scale;
  float tn[kTnPre];
};

// Window buffer b of the sweep's LDS by address arithmetic (the buffers are
// contiguous), so that no array of pointers is indexed at run time.
template <typename L_>
__device__ __forceinline__ Win pick_win(const L_& L, int b) {
  Win w;
  w.base = L.win[0].base + b * L.wstride;
  w.av = L.win[0].av + b * L.wstride;
  return w;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

template <int kW, bool kSmall, bool kHalf = false>
__device__ __forceinline__ void awin_fetch(const SwGeom& G, const ColumnArgs& A, int c, int s0,
                                           int nv, APre<kSmall, kHalf>& pre) {
  const int lane = lane_id();
  const size_t base = (size_t)c * G.nslots + s0;
#pragma unroll
  for (int i = 0; i < APre<kSmall, kHalf>::PI; ++i) {
    const int j = lane + i * kWave;
    pre.a0[i] = j < nv ? A.afirst[base + j] : -1;
  }
  const int nav = nv * kW;
#pragma unroll
  for (int i = 0; i < APre<kSmall, kHalf>::PA; ++i) {
    const int j = lane + i * kWave;
    if constexpr (kHalf)
      pre.av[i] = j < nav ? A.aval16[base * kW + j] : (unsigned short)0;
    else
      pre.av[i] = j < nav ? A.aval[base * kW + j] : 0.f;
  }
  if constexpr (kHalf) pre.scale = A.ascale[c];
}

// The t-1 and t+1 columns of voxel column c (not updated in a sweep launch).
template <typename Pre>
__device__ __forceinline__ void tn_fetch(const SwGeom& G, const float* x, int f, int c,
                                         size_t vol, Pre& pre) {
#pragma unroll
  for (int i = 0; i < kTnPre; ++i) {
    const int j = lane_id() + i * kWave;
    const int ff = j < G.N_z ? f - 1 : f + 1;
    const bool ok = j < 2 * G.N_z && ff >= 0 && ff < G.N_time;
    pre.tn[i] = ok ? x[(size_t)ff * vol + (size_t)c * G.N_z + j % G.N_z] : 0.f;
  }
}

// Stores a prefetched window (first bins turned into sinogram offsets
// through the frame's view list) and t+-1 values.
template <int kW, bool kSmall, bool kHalf = false>
__device__ __forceinline__ void awin_store(const SwGeom& G, const APre<kSmall, kHalf>& pre,
                                           int nv, const int* views, const Win& win, float* tn) {
  const int lane = lane_id();
#pragma unroll
  for (int i = 0; i < APre<kSmall, kHalf>::PI; ++i) {
    const int j = lane + i * kWave;
    if (j < nv) win.base[j] = pre.a0[i] >= 0 ? (views[j] * G.N_r + pre.a0[i]) * G.N_t : -1;
  }
  const int nav = nv * kW;
#pragma unroll
  for (int i = 0; i < APre<kSmall, kHalf>::PA; ++i) {
    const int j = lane + i * kWave;
    if constexpr (kHalf) {
      // the same product k_aquant wrote into the FP32 copy of A
      if (j < nav) win.av[j] = (float)__builtin_bit_cast(_Float16, pre.av[i]) * pre.scale;
    } else {
      if (j < nav) win.av[j] = pre.av[i];
    }
  }
#pragma unroll
  for (int i = 0; i < kTnPre; ++i) {
    const int j = lane + i * kWave;
    if (j < 2 * G.N_z) tn[j] = pre.tn[i];
  }
}

// Lane layout over a column's sinogram footprint: lane = (detector row rr,
// slot group g), g fastest; a lane takes view slots g, g + groups, ... in row
// rr (and rr + rpl, ... when N_t > 64).  A slot's W bins lie N_t apart in e
// and w.  With a power-of-two rpl the slot groups of a row are `groups`
// consecutive lanes, summed with DPP.
struct Lanes {
  int rpl, groups, g, rr;
  bool pow2;
  __device__ Lanes(int N_t) {
    const int lane = lane_id();
    rpl = N_t < kWave ? N_t : kWave;
    groups = kWave / rpl;
    rr = lane / groups;
    g = lane % groups;
    pow2 = (rpl & (rpl - 1)) == 0;
  }
};

// theta1/theta2 per detector row of one voxel column (XT_VoxUpdate.c:56-85,
// generalized-Huber weight included).  Each lane issues the e/w loads of a
// batch of its slots back to back, then accumulates; row sums land in
// row1/row2 (part1/part2: LDS partials for the non-power-of-two case).
template <int kW>
__device__ __forceinline__ void col_thetas(const SwGeom& G, const ColumnArgs& A, const Win& win,
                                           int nv, float* row1, float* row2, float* part1,
                                           float* part2) {
  constexpr int SB = kW <= 5 ? 2 : 1;  // slots per load batch
  const int lane = lane_id();
  const Lanes L(G.N_t);
  const float T2 = G.huber_T * G.huber_T, Thd = G.huber_T * G.huber_delta;
  for (int t0 = 0; t0 < G.N_t; t0 += L.rpl) {
    const int row = t0 + L.rr;
    const bool lane_ok = L.rr < L.rpl && row < G.N_t;
    float a1 = 0.f, a2 = 0.f;
    for (int sb = 0; sb < nv; sb += SB * L.groups) {
      float ev[SB][kW], wv[SB][kW], av[SB][kW];
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int sl = sb + L.g + b * L.groups;
        const int base = (lane_ok && sl < nv) ? win.base[sl] : -1;
#pragma unroll
        for (int k = 0; k < kW; ++k) {
          const float a = base >= 0 ? win.av[sl * kW + k] : 0.f;
          const bool ok = a != 0.f;
          const int o = ok ? base + k * G.N_t + row : 0;
          const float e = e_load(A.e + o);
          ev[b][k] = ok ? e : 0.f;
          wv[b][k] = ok ? A.w[o] : 0.f;
          av[b][k] = a * G.delta_t;
        }
      }
#pragma unroll
      for (int b = 0; b < SB; ++b)
#pragma unroll
        for (int k = 0; k < kW; ++k) {
          // |e sqrt(w)| < T tested as e^2 w < T^2; the generalized-Huber
          // weight T delta sqrt(w)/|e| only where an entry is an outlier
          const float e = ev[b][k], w = wv[b][k], a = av[b][k];
          float hw = w;
          if (e * e * w >= T2) hw = Thd * sqrtf(w) * __builtin_amdgcn_rcpf(fabsf(e));
          a2 += a * a * hw;
          a1 -= e * a * hw;
        }
    }
    if (L.pow2) {
      a1 = group_sum(a1, L.groups);
      a2 = group_sum(a2, L.groups);
      if (L.g == 0 && row < G.N_t) {
        row1[row] = a1;
        row2[row] = a2;
      }
    } else {
      part1[lane] = a1;
      part2[lane] = a2;
      wave_sync();
      if (lane < L.rpl && t0 + lane < G.N_t) {
        float s1 = 0.f, s2 = 0.f;
        for (int gg = 0; gg < L.groups; ++gg) {
          s1 += part1[lane * L.groups + gg];
          s2 += part2[lane * L.groups + gg];
        }
        row1[t0 + lane] = s1;
        row2[t0 + lane] = s2;
      }
    }
    wave_sync();
  }
}

// Pipelined theta (the column's window fits one load batch: every slot of
// a lane's row in SB slots, N_t <= 64 and a power of two): the e/w loads of
// the next column go out while the current column is solved, and are
// consumed after it, corrected for the current column's own error update,
// which reaches the shared sinogram only after them.
template <int kW>
struct ThetaLoads {
  static constexpr int SB = kW <= 5 ? 2 : 1;
  float ev[SB][kW], wv[SB][kW];
};

template <int kW>
__device__ __forceinline__ void theta_issue(const SwGeom& G, const ColumnArgs& A, const Win& win,
                                            int nv, ThetaLoads<kW>& tl) {
  constexpr int SB = ThetaLoads<kW>::SB;
  const Lanes L(G.N_t);
  const int row = L.rr;
  const bool lane_ok = L.rr < L.rpl && row < G.N_t;
#pragma unroll
  for (int b = 0; b < SB; ++b) {
    const int sl = L.g + b * L.groups;
    const int base = (lane_ok && sl < nv) ? win.base[sl] : -1;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
      const float a = base >= 0 ? win.av[sl * kW + k] : 0.f;
      const bool ok = a != 0.f;
      const int o = ok ? base + k * G.N_t + row : 0;
      // unconditional loads (a fixed count per lane keeps the compiler's
      // wait counts exact); entries with a = 0 are not used
      tl.ev[b][k] = e_load(A.e + o);
      tl.wv[b][k] = A.w[o];
    }
  }
}

// theta1/theta2 per row from loads issued by theta_issue on window `win`,
// first subtracting the previous column's own update (window `wprev`, voxel
// changes dxp; has_prev = 0 for the first column), as the sinogram held no
// part of it when the loads went out.
template <int kW>
__device__ __forceinline__ void theta_finish(const SwGeom& G, const Win& win, const Win& wprev,
                                             bool has_prev, const float* dxp, int lognt,
                                             int nv, const ThetaLoads<kW>& tl, float* row1,
                                             float* row2) {
  constexpr int SB = ThetaLoads<kW>::SB;
  const Lanes L(G.N_t);
  const int row = L.rr;
  const bool lane_ok = L.rr < L.rpl && row < G.N_t;
  const float T2 = G.huber_T * G.huber_T, Thd = G.huber_T * G.huber_delta;
  const float dzp = (has_prev && lane_ok) ? dxp[row / G.zo] : 0.f;
  float a1 = 0.f, a2 = 0.f;
#pragma unroll
  for (int b = 0; b < SB; ++b) {
    const int sl = L.g + b * L.groups;
    const bool sok = lane_ok && sl < nv;
    const int base = sok ? win.base[sl] : -1;
    const int bprev = (sok && dzp != 0.f) ? wprev.base[sl] : -1;
    // bin offset of this window's first bin in the previous window
    const int shift = (base >= 0 && bprev >= 0) ? (base - bprev) >> lognt : kW;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
      const float a = base >= 0 ? win.av[sl * kW + k] : 0.f;
      const bool ok = a != 0.f;
      float e = ok ? tl.ev[b][k] : 0.f;
      const int d = shift + k;
      if (ok && d >= 0 && d < kW) e -= (wprev.av[sl * kW + d] * G.delta_t) * dzp;
      const float w = ok ? tl.wv[b][k] : 0.f, av = a * G.delta_t;
      float hw = w;
      if (e * e * w >= T2) hw = Thd * sqrtf(w) * __builtin_amdgcn_rcpf(fabsf(e));
      a2 += av * av * hw;
      a1 -= e * av * hw;
    }
  }
  a1 = group_sum(a1, L.groups);
  a2 = group_sum(a2, L.groups);
  if (L.g == 0 && lane_ok) {
    row1[row] = a1;
    row2[row] = a2;
  }
}

// e -= A delta_t dx along one voxel column (XT_VoxUpdate.c:96-116), with
// atomics because other wavefronts' columns share detector bins.
template <int kW>
__device__ __forceinline__ void col_update(const SwGeom& G, const ColumnArgs& A, const Win& win,
                                           int nv, const float* dxz) {
  const Lanes L(G.N_t);
  for (int t0 = 0; t0 < G.N_t; t0 += L.rpl) {
    const int row = t0 + L.rr;
    if (L.rr >= L.rpl || row >= G.N_t) continue;
    const float dz = dxz[row / G.zo];
    if (dz == 0.f) continue;
    for (int sl = L.g; sl < nv; sl += L.groups) {
      const int base = win.base[sl];
      if (base < 0) continue;
#pragma unroll
      for (int k = 0; k < kW; ++k) {
        const float a = win.av[sl * kW + k];
        if (a != 0.f) e_add(A.e + base + k * G.N_t + row, -(a * G.delta_t) * dz);
      }
    }
  }
}

// Lanes per voxel for the fixed part of the surrogate: the largest power of
// two with lanes-per-voxel x min(N_z, 64) <= 64.
__device__ __forceinline__ int fixed_lpv(int Nz) {
  const int v = Nz < kWave ? Nz : kWave;
  int lpv = 1;
  while (2 * lpv * v <= kWave) lpv <<= 1;
  return lpv;
}

// Neighbour m (0..25) outside the voxel's own column: the 24 spatial
// neighbours ([dz][dy][dx] index j, skipping 4, 13 and 22, the column's own
// z-1, centre and z+1) and t-1 / t+1 (j = 27, 28).
__device__ __forceinline__ int fixed_nb(int m) {
  return m < 24 ? m + (m >= 4) + (m >= 12) + (m >= 20) : m + 3;
}

// The part of each voxel's surrogate (XT_ICD_update.c:133-182) that the
// column's own update leaves unchanged: the terms of its neighbours outside
// the column, evaluated once per column at the voxel's current value V.
// lpv consecutive lanes per voxel, summed with DPP; writes fnum (sum of
// weight x neighbour), fden (sum of weights) and fpos (some neighbour > 0).
__device__ __forceinline__ void col_fixed(const SwGeom& G, const PriorTab& P, const float* sx,
                                          int BW, int ly, int lx, int cy, int cx, int f,
                                          const float* tn, float* fnum, float* fden,
                                          int* fpos) {
  const int lane = lane_id(), Nz = G.N_z;
  const int lpv = fixed_lpv(Nz), vpp = kWave / lpv;
  const int sub = lane % lpv;
  for (int zb = 0; zb < Nz; zb += vpp) {
    const int z = zb + lane / lpv;
    const bool act = z < Nz;
    const int centre = (ly * BW + lx) * Nz + z;
    const float V = act ? sx[centre] : 0.f;
    float num = 0.f, den = 0.f;
    int pos = 0;
    if (act) {
      for (int m = sub; m < 26; m += lpv) {
        const int j = fixed_nb(m);
        bool ok;
        float nbv;
        if (j < 27) {
          const int dz = j / 9 - 1, dy = (j / 3) % 3 - 1, dx = j % 3 - 1;
          ok = z + dz >= 0 && z + dz < Nz && cy + dy >= 0 && cy + dy < G.N_y && cx + dx >= 0 &&
               cx + dx < G.N_x;
          nbv = ok ? sx[centre + (dy * BW + dx) * Nz + dz] : 0.f;
        } else {
          const int ff = j == 27 ? f - 1 : f + 1;
          ok = ff >= 0 && ff < G.N_time;
          nbv = ok ? tn[(j - 27) * Nz + z] : 0.f;
        }
        if (nbv > 0.f) pos = 1;
        if (ok) {
          const float wt = nb_weight(P, j, V, nbv);
          num += wt * nbv;
          den += wt;
        }
      }
    }
    num = group_sum(num, lpv);
    den = group_sum(den, lpv);
    pos = group_or(pos, lpv);
    if (act && sub == 0) {
      fnum[z] = num;
      fden[z] = den;
      fpos[z] = pos;
    }
  }
}

// Prior constants in registers (the prior wavefront loads them once).
struct PriorReg {
  float qmp[2], c[2], inv_sq[2], inv_sqp[2], b0[2];
  float f4, f22;  // weights of the column's own z-1 / z+1 neighbours
  __device__ void load(const PriorTab& P) {
    for (int t = 0; t < 2; ++t) {
      qmp[t] = P.qmp[t];
      c[t] = P.c[t];
      inv_sq[t] = P.inv_sq[t];
      inv_sqp[t] = P.inv_sqp[t];
      b0[t] = P.b0[t];
    }
    f4 = P.filt[4];
    f22 = P.filt[22];
  }
  // b(d) of prior t (0 spatial, 1 temporal)
  __device__ __forceinline__ float b(bool temporal, float d) const {
    return temporal ? qg_b(d, qmp[1], c[1], inv_sq[1], inv_sqp[1], b0[1])
                    : qg_b(d, qmp[0], c[0], inv_sq[0], inv_sqp[0], b0[0]);
  }
};

// A lane's share of the out-of-column neighbours of its voxel, fixed for
// the whole sweep when lpv >= 8 (N_z <= 8: at most 4 of the 26 per lane):
// LDS offset (from the voxel centre in the slab, or from the voxel's z in
// the t+-1 buffer), weight, and validity: a z / temporal condition fixed for
// the lane, and the tile-edge conditions (bits: needs row above / below,
// column left / right) tested per column.
constexpr int kNbLane = 4;
struct NbTab {
  int off[kNbLane];
  float wt[kNbLane];
  unsigned needs[kNbLane];  // bits 0-3 edges; bit 4 temporal; bit 5 usable
};

__device__ __forceinline__ void nbtab_init(const SwGeom& G, const PriorTab& P, int BW, int f,
                                           NbTab& T) {
  const int lane = lane_id(), Nz = G.N_z;
  const int lpv = fixed_lpv(Nz), sub = lane % lpv, z = lane / lpv;
#pragma unroll
  for (int i = 0; i < kNbLane; ++i) {
    const int m = sub + i * lpv;
    T.off[i] = 0;
    T.wt[i] = 0.f;
    T.needs[i] = 0;
    if (m >= 26 || z >= Nz) continue;
    const int j = fixed_nb(m);
    if (j < 27) {
      const int dz = j / 9 - 1, dy = (j / 3) % 3 - 1, dx = j % 3 - 1;
      if (z + dz < 0 || z + dz >= Nz) continue;
      T.off[i] = (dy * BW + dx) * Nz + dz;
      T.wt[i] = P.filt[j];
      T.needs[i] = (dy < 0 ? 1u : 0u) | (dy > 0 ? 2u : 0u) | (dx < 0 ? 4u : 0u) |
                   (dx > 0 ? 8u : 0u) | 32u;
    } else {
      const int ff = j == 27 ? f - 1 : f + 1;
      if (ff < 0 || ff >= G.N_time) continue;
      T.off[i] = (j - 27) * Nz;
      T.wt[i] = P.filt[27];
      T.needs[i] = 16u | 32u;
    }
  }
}

// col_fixed for lpv >= 8 from the lane's table: per neighbour one edge test,
// one LDS read and one qGGMRF weight.
__device__ __forceinline__ void col_fixed_tab(const SwGeom& G, const PriorReg& R, const NbTab& T,
                                              const float* sx, int BW, int ly, int lx, int cy,
                                              int cx, const float* tn, float* fnum, float* fden,
                                              int* fpos) {
  const int lane = lane_id(), Nz = G.N_z;
  const int lpv = fixed_lpv(Nz), sub = lane % lpv, z = lane / lpv;
  const bool act = z < Nz;
  const int centre = (ly * BW + lx) * Nz + z;
  const unsigned edge = (cy == 0 ? 1u : 0u) | (cy == G.N_y - 1 ? 2u : 0u) |
                        (cx == 0 ? 4u : 0u) | (cx == G.N_x - 1 ? 8u : 0u);
  const float V = act ? sx[centre] : 0.f;
  float nbv[kNbLane];
  bool ok[kNbLane];
#pragma unroll
  for (int i = 0; i < kNbLane; ++i) {
    const unsigned nd = T.needs[i];
    ok[i] = act && (nd & 32u) && !(nd & edge & 15u);
    const float* src = (nd & 16u) ? tn + z : sx + centre;
    nbv[i] = ok[i] ? src[T.off[i]] : 0.f;
  }
  float num = 0.f, den = 0.f;
  int pos = 0;
#pragma unroll
  for (int i = 0; i < kNbLane; ++i) {
    if (nbv[i] > 0.f) pos = 1;
    const float wt = ok[i] ? T.wt[i] * R.b((T.needs[i] & 16u) != 0, V - nbv[i]) : 0.f;
    num += wt * nbv[i];
    den += wt;
  }
  num = group_sum(num, lpv);
  den = group_sum(den, lpv);
  pos = group_or(pos, lpv);
  if (act && sub == 0) {
    fnum[z] = num;
    fden[z] = den;
    fpos[z] = pos;
  }
}

struct SolveStats {
  double dx = 0.0, x = 0.0, skip = 0.0, upd = 0.0;
};

// Solves the voxels of one z parity of the column, one lane per voxel: the
// fixed part plus the column's own z-1 / z+1 neighbours at their current
// values, zero skipping (XT_VoxUpdate.c:238-262), the surrogate minimiser and
// over-relaxation (XT_ICD_update.c:174-176).  New values go to the slab,
// changes to dxz.
__device__ __forceinline__ void col_solve(const SwGeom& G, const PriorReg& R, float* sx,
                                          int centre0, int par, int iter, float alpha,
                                          const float* fnum, const float* fden,
                                          const int* fpos, const float* row1,
                                          const float* row2, float* dxz, SolveStats& st) {
  const int Nz = G.N_z;
  const int nzp = (Nz - par + 1) / 2;
  const int vi = lane_id();
  if (vi >= nzp) return;
  const int z = par + 2 * vi;
  const int centre = centre0 + z;
  const float V = sx[centre];
  float num = fnum[z], den = fden[z];
  int pos = fpos[z];
  if (z > 0) {
    const float nbv = sx[centre - 1];
    const float wt = R.f4 * R.b(false, V - nbv);
    num += wt * nbv;
    den += wt;
    if (nbv > 0.f) pos = 1;
  }
  if (z + 1 < Nz) {
    const float nbv = sx[centre + 1];
    const float wt = R.f22 * R.b(false, V - nbv);
    num += wt * nbv;
    den += wt;
    if (nbv > 0.f) pos = 1;
  }
  const bool skip = V == 0.f && iter > 1 && !pos;
  float xn = V;
  if (!skip) {
    float th1 = 0.f, th2 = 0.f;
    for (int m = 0; m < G.zo; ++m) {
      th1 += row1[z * G.zo + m];
      th2 += row2[z * G.zo + m];
    }
    xn = substitute(alpha, V, th1, th2, num, den);
  }
  sx[centre] = xn;
  dxz[z] = xn - V;
  st.skip += skip ? 1.0 : 0.0;
  st.upd += skip ? 0.0 : 1.0;
  st.dx += fabs((double)xn - (double)V);
  st.x += skip ? 0.0 : fabs((double)xn);
}

struct SweepArgs {
  ColumnArgs col;
  const int* frame_ptr;
  const int* tile_ptr;
  const int* tile_cols;
  const int* launch_tiles;
  int n_launch_tiles;
  // paired launches (option pair_parities): workgroups from nwg1 on run a
  // second (colour, frame parity) set, its tiles, count, parity and seed
  int nwg1;
  const int* launch_tiles2;
  int n_launch_tiles2, fparity2;
  unsigned seed2;
  int tiles_x;  // tiles per row of the tile grid
  int tile;     // tile edge T
  int nvmax;    // most views in one frame (sizes the LDS windows)
  int fparity;
  int iter;            // iteration number (offset from *iter_dev when set)
  const int* iter_dev;  // graph replays: the first iteration, in device memory
  unsigned seed;
  float alpha;
  float* x;
  double* stats;
  // non-homogeneous ICD: mag, when set, receives each visited column's sum of
  // |dx| ([N_time][N_y N_x]); nh_frac > 0 makes each workgroup visit only the
  // floor(nh_frac n) (at least one) of its tile's n columns that rank highest
  // in nh_rank (same layout), ties to the earlier list entry
  float* mag;
  const float* nh_rank;
  float nh_frac;
  // placement of the launch's (tile, frame) items on the XCDs (workgroup w
  // runs on XCD w mod 8 when the dispatcher deals workgroups round-robin):
  // place_r = 0: tile-major grid (w = frame x n_launch_tiles + tile);
  // place_r > 0: each of the nfl frames split into place_r parts (tiles
  // t = part mod place_r), unit u = part x nfl + frame on XCD u mod 8, the
  // unit's items in unit_slots consecutive slots of that XCD
  int place_r, nfl, unit_slots;
  unsigned* xcd_check;  // [0] items run, [1] items off their XCD (may be null)
};

constexpr int kSweepThreads = 2 * kWave;
constexpr int kXcds = 8;          // XCDs of an MI355X, each with its own L2
constexpr int kHwRegXccId = 20;   // s_getreg id of XCC_ID (gfx940 family)
constexpr int kNhMax = 2 * kSweepThreads;  // tile columns a non-homogeneous pass can rank

// One ICD pass over the voxel columns of the tiles in this launch (one
// 2x2-checkerboard colour) for the frames of one parity, one workgroup of two
// wavefronts per (tile, frame).  The tile's object slab plus a one-column
// halo lives in LDS for the whole pass (neighbouring tiles and adjacent frames
// are not updated in the same launch, so the halo and the t+-1 values are
// constant), with the tile's column list and the frame's view list.  Per
// column, in two barrier-separated phases:
//   1. wavefront 0 computes theta1/theta2 from the shared error sinogram and
//      prefetches the next column's A window and t+-1 values into registers,
//      while wavefront 1 evaluates the surrogate's terms of the neighbours
//      outside the column;
//   2. wavefront 1 solves the z-parity halves (each voxel sees its z+-1
//      neighbours' new values) while wavefront 0 stores the prefetch;
// then wavefront 0 sends the column's error update with atomics and moves on
// to the next column's theta while wavefront 1 starts its terms.
// kW = W, the A window width.
// The small-frame variants with window widths up to 5 are held to 128 VGPRs
// (four wavefronts per SIMD, two spilled registers at W = 5): at 128x128x4
// with tile 4 a launch's 2048 workgroups then fit the chip at once instead
// of in two rounds.
template <int kW, bool kSmall, bool kHalf>
__global__ __launch_bounds__(kSweepThreads)
__attribute__((amdgpu_waves_per_eu((kSmall && kW <= 5) ? 4 : 1))) void k_sweep(SwGeom G,
                                                                              SweepArgs A) {
  extern __shared__ __align__(16) float smem[];
  __shared__ float s_tn[kWinBufs][2 * kNZMax];
  __shared__ float s_row1[kNTMax], s_row2[kNTMax];
  __shared__ float s_part1[kWave], s_part2[kWave];
  __shared__ float s_dx[kNZMax], s_fnum[kNZMax], s_fden[kNZMax];
  __shared__ int s_fpos[kNZMax];
  __shared__ PriorTab s_prior;
  const int tid = threadIdx.x;
  // wavefront 0: sinogram side; 1: prior side.  Read from lane 0 into a
  // scalar register, so the role branch is a uniform (scalar) branch and each
  // wavefront executes only its own loop's barriers.
  const bool w0 = __builtin_amdgcn_readfirstlane(tid) < kWave;
  const int wid_all = blockIdx.x;
  const int iter = A.iter_dev ? *A.iter_dev + A.iter : A.iter;
  load_prior(G, &s_prior);
  // this workgroup's (colour, frame parity) set
  int wid = wid_all, ntl = A.n_launch_tiles, fpar = A.fparity;
  const int* ltiles = A.launch_tiles;
  unsigned lseed = A.seed;
  if (wid_all >= A.nwg1) {
    wid = wid_all - A.nwg1;
    ntl = A.n_launch_tiles2;
    fpar = A.fparity2;
    ltiles = A.launch_tiles2;
    lseed = A.seed2;
  }
  int tl, fl;
  if (A.place_r == 0) {
    tl = wid % ntl;
    fl = wid / ntl;
  } else {
    const int s = wid / kXcds, u = wid % kXcds + kXcds * (s / A.unit_slots);
    fl = u % A.nfl;
    tl = u / A.nfl + A.place_r * (s % A.unit_slots);
    if (tl >= ntl) return;
  }
  const int tile = ltiles[tl];
  const int f = fpar + 2 * fl;
  if (f >= G.N_time) return;
  if (A.xcd_check && tid == 0) {
    // the XCD this workgroup runs on (hardware register XCC_ID, bits 3:0)
    const unsigned xcc = (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | kHwRegXccId) & 15u;
    atomicAdd(&A.xcd_check[0], 1u);
    if (xcc != (unsigned)(wid_all % kXcds)) atomicAdd(&A.xcd_check[1], 1u);
  }
  const int c0 = A.tile_ptr[tile];
  const int n_all = A.tile_ptr[tile + 1] - c0;
  int n = n_all;
  if (A.nh_frac > 0.f) {
    n = (int)(A.nh_frac * (float)n_all);
    if (n < 1) n = 1;
    if (n > n_all) n = n_all;
  }
  const int s0 = A.frame_ptr[f];
  const int nv = A.frame_ptr[f + 1] - s0;
  if (n == 0 || nv == 0) return;
  const int Nz = G.N_z, T = A.tile, BW = T + 2;
  const int nx = BW * BW * Nz;
  const SweepLds L(smem, nx, A.nvmax, kW);
  const int oy = (tile / A.tiles_x) * T - 1, ox = (tile % A.tiles_x) * T - 1;
  const size_t vol = (size_t)G.N_y * G.N_x * Nz;
  float* xf = A.x + (size_t)f * vol;
  // seeded random visiting order: k -> (pa k + pb) mod n, gcd(pa, n) = 1
  const unsigned h = mix32(lseed ^ mix32((unsigned)tile * 131071u + (unsigned)f * 524287u));
  int pa = 1, pb = 0;
  if (n > 1) {
    pa = 1 + (int)(h % (unsigned)(n - 1));
    while (gcd_i(pa, n) != 1) pa = pa % (n - 1) + 1;
    pb = (int)(mix32(h + 0x9e3779b9u) % (unsigned)n);
  }
  // the tile's slab with halo (8 loads per thread in flight), column and view lists
  for (int j0 = tid; j0 < nx; j0 += 8 * kSweepThreads) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = j0 + u * kSweepThreads;
      const int z = j % Nz, b = j / Nz;
      const int yy = oy + b / BW, xx = ox + b % BW;
      v[u] = (j < nx && yy >= 0 && yy < G.N_y && xx >= 0 && xx < G.N_x)
                 ? xf[((size_t)yy * G.N_x + xx) * Nz + z]
                 : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u * kSweepThreads < nx) L.x[j0 + u * kSweepThreads] = v[u];
  }
  for (int j = tid; j < n_all; j += kSweepThreads) L.cols[j] = A.tile_cols[c0 + j];
  for (int j = tid; j < nv; j += kSweepThreads) L.views[j] = A.col.frame_views[s0 + j];
  if (n < n_all) {
    // non-homogeneous pass: rank the tile's columns by nh_rank (the count of
    // columns ranked above, ties to the earlier entry, so the ranks are a
    // permutation) and compact the n highest to the front of the list
    const float* rk = A.nh_rank + (size_t)f * G.N_y * G.N_x;
    for (int j = tid; j < n_all; j += kSweepThreads) s_row1[j] = rk[L.cols[j]];
    __syncthreads();
    int keep_c[kNhMax / kSweepThreads], keep_r[kNhMax / kSweepThreads];
#pragma unroll
    for (int u = 0; u < kNhMax / kSweepThreads; ++u) {
      const int j = tid + u * kSweepThreads;
      keep_r[u] = n_all;
      keep_c[u] = 0;
      if (j < n_all) {
        const float v = s_row1[j];
        int r = 0;
        for (int i = 0; i < n_all; ++i) {
          const float o = s_row1[i];
          r += (o > v || (o == v && i < j)) ? 1 : 0;
        }
        keep_r[u] = r;
        keep_c[u] = L.cols[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kNhMax / kSweepThreads; ++u)
      if (keep_r[u] < n) L.cols[keep_r[u]] = keep_c[u];
  }
  __syncthreads();
  int pos = pb;  // visiting order: pos_k = (pb + k pa) mod n
  // the theta pipeline applies when a column's window is one load batch per
  // lane (every slot of a lane's row in SB slots, N_t <= 64, a power of two)
  constexpr int SB = ThetaLoads<kW>::SB;
  const int rpl = G.N_t < kWave ? G.N_t : kWave;
  const bool pipe = G.N_t <= kWave && (G.N_t & (G.N_t - 1)) == 0 && nv <= SB * (kWave / rpl);
  if (w0) {
    // window and t+-1 values of the first column
    APre<kSmall, kHalf> pre;
    const int cf = L.cols[pos];
    awin_fetch<kW, kSmall, kHalf>(G, A.col, cf, s0, nv, pre);
    tn_fetch(G, A.x, f, cf, vol, pre);
    awin_store<kW, kSmall, kHalf>(G, pre, nv, L.views, L.win[0], s_tn[0]);
  }
  // nothing stays in flight: keeps the compiler's wait analysis from
  // carrying these loads into the column loop
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  SolveStats st;
  // Each wavefront runs its own loop over the columns; both loops pass the
  // same two workgroup barriers per column, so the phases stay aligned.
  if (w0) {
    const int lognt = __builtin_ctz((unsigned)G.N_t);
    ThetaLoads<kW> tl;
    // pipeline: the prefetch of column k+1 is in flight during column k's
    // solve, and stored at the start of column k's step (before the theta
    // loads of column k+1 go out)
    APre<kSmall, kHalf> pre;
    if (pipe) {
      if (n > 1) {
        int p1 = pos + pa;
        if (p1 >= n) p1 -= n;
        const int cn = L.cols[p1];
        awin_fetch<kW, kSmall, kHalf>(G, A.col, cn, s0, nv, pre);
        tn_fetch(G, A.x, f, cn, vol, pre);
      }
      theta_issue<kW>(G, A.col, L.win[0], nv, tl);
    }
    for (int k = 0; k < n; ++k) {
      const int bc = k % kWinBufs, bn = (k + 1) % kWinBufs, bp = (k + 2) % kWinBufs;
      const Win wcur = pick_win(L, bc), wnext = pick_win(L, bn);
      const Win wprev = pick_win(L, bp);  // column k-1's window
      int p1 = pos + pa;
      if (p1 >= n) p1 -= n;
      int p2 = p1 + pa;
      if (p2 >= n) p2 -= n;
      pos = p1;
      if (pipe) {
        // phase 1: theta of column k from the loads issued during the last
        // solve, corrected for column k-1's own update; that update's
        // atomics; column k+1's prefetch (it landed with the theta loads)
        // stored over column k-2's window; the prefetch of column k+2
        theta_finish<kW>(G, wcur, wprev, k > 0, s_dx, lognt, nv, tl, s_row1, s_row2);
        if (k > 0) col_update<kW>(G, A.col, wprev, nv, s_dx);
        if (k + 1 < n) awin_store<kW, kSmall, kHalf>(G, pre, nv, L.views, wnext, s_tn[bn]);
        if (k + 2 < n) {
          const int cn = L.cols[p2];
          awin_fetch<kW, kSmall, kHalf>(G, A.col, cn, s0, nv, pre);
          tn_fetch(G, A.x, f, cn, vol, pre);
        }
        __syncthreads();
        // phase 2 (during the solve of column k): column k+1's theta loads
        if (k + 1 < n) theta_issue<kW>(G, A.col, wnext, nv, tl);
        __syncthreads();
      } else {
        // phase 1: prefetch of column k+1, theta of column k
        if (k + 1 < n) {
          const int cn = L.cols[p1];
          awin_fetch<kW, kSmall, kHalf>(G, A.col, cn, s0, nv, pre);
          tn_fetch(G, A.x, f, cn, vol, pre);
        }
        col_thetas<kW>(G, A.col, wcur, nv, s_row1, s_row2, s_part1, s_part2);
        __syncthreads();
        // phase 2: store the prefetch (it landed with the theta loads) and
        // retire every load explicitly (a free wait here) so that the
        // compiler's wait analysis carries none into the next column, before
        // the error update's atomics go out
        if (k + 1 < n) awin_store<kW, kSmall, kHalf>(G, pre, nv, L.views, wnext, s_tn[bn]);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        col_update<kW>(G, A.col, wcur, nv, s_dx);
      }
    }
    if (pipe) col_update<kW>(G, A.col, pick_win(L, (n - 1) % kWinBufs), nv, s_dx);
  } else {
    PriorReg R;
    R.load(s_prior);
    NbTab NT;
    const bool tab = fixed_lpv(Nz) >= 8;
    if (tab) nbtab_init(G, s_prior, BW, f, NT);
    for (int k = 0; k < n; ++k) {
      const float* tcur = s_tn[k % kWinBufs];
      const int c = L.cols[pos];
      pos += pa;
      if (pos >= n) pos -= n;
      const int cy = c / G.N_x, cx = c % G.N_x;
      const int ly = cy - oy, lx = cx - ox;
      const int centre0 = (ly * BW + lx) * Nz;
      // phase 1: the surrogate's terms of the neighbours outside the column
      if (tab)
        col_fixed_tab(G, R, NT, L.x, BW, ly, lx, cy, cx, tcur, s_fnum, s_fden, s_fpos);
      else
        col_fixed(G, s_prior, L.x, BW, ly, lx, cy, cx, f, tcur, s_fnum, s_fden, s_fpos);
      __syncthreads();
      // phase 2: the z-parity halves
      col_solve(G, R, L.x, centre0, 0, iter, A.alpha, s_fnum, s_fden, s_fpos, s_row1, s_row2,
                s_dx, st);
      wave_sync();
      col_solve(G, R, L.x, centre0, 1, iter, A.alpha, s_fnum, s_fden, s_fpos, s_row1, s_row2,
                s_dx, st);
      __syncthreads();
      if (A.mag) {
        // the column's sum of |dx| (s_dx holds it until the next solve)
        float m = 0.f;
        for (int z = lane_id(); z < Nz; z += kWave) m += fabsf(s_dx[z]);
        for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off);
        if (lane_id() == 0) A.mag[(size_t)f * G.N_y * G.N_x + c] = m;
      }
    }
  }
  __syncthreads();
  // write the tile back (interior of the box, inside the object)
  for (int j = tid; j < nx; j += kSweepThreads) {
    const int z = j % Nz, b = j / Nz;
    const int by = b / BW, bx = b % BW;
    if (by == 0 || bx == 0 || by == BW - 1 || bx == BW - 1) continue;
    const int yy = oy + by, xx = ox + bx;
    if (yy < G.N_y && xx < G.N_x) xf[((size_t)yy * G.N_x + xx) * Nz + z] = L.x[j];
  }
  if (!w0) {
    for (int off = 32; off > 0; off >>= 1) {
      st.dx += __shfl_xor(st.dx, off);
      st.x += __shfl_xor(st.x, off);
      st.skip += __shfl_xor(st.skip, off);
      st.upd += __shfl_xor(st.upd, off);
    }
    if (tid == kWave) {
      atomicAdd(&A.stats[0], st.dx);
      atomicAdd(&A.stats[1], st.x);
      atomicAdd(&A.stats[2], st.skip);
      atomicAdd(&A.stats[3], st.upd);
    }
  }
}

// ----------------------------------------------------------------------
// PCMBIR tomographic ICD (the ENABLE_TOMO_RECONS path of PCMBIR_3D,
// compiled out of the shipped reference): two objects, delta and beta, on
// the A rule times the wavenumber (PCMBIR_3D/XT_AMatrix.c:103), each with its
// error sinogram, updated together by the coupled substitution of
// PCMBIR_3D/XT_Prior.c:137-239.  Index 0 of the paired constants is the
// phase (delta) prior, 1 the magnitude (beta) prior.
struct PcTab {
  float T[4];                    // DecorrTran, row-major
  float qmp;                     // q - p
  float c[2][2], inv_sq[2][2], inv_sqp[2][2], b0[2][2];  // [spatial, temporal][phase, mag]
  float filt[28];                // 27 spatial ([dz][dy][dx]) + temporal
  float mu, kw, alpha;
};

struct PcArgs {
  const float* aval;
  const int* afirst;
  const int* frame_views;
  const int* frame_ptr;
  const int* tile_ptr;
  const int* tile_cols;
  const int* launch_tiles;
  int n_launch_tiles, tiles_x, tile, nvmax, fparity, iter, W;
  unsigned seed;
  float *xd, *xb, *ed, *eb;
  double* stats;  // [sum |dx|, updated, skipped]
};

// Adds one neighbour's share of the coupled surrogate: D = V - n, the
// decorrelated differences T0 D (phase prior) and T1 D (magnitude prior)
// give b0, b1, and Q = T^T diag(b0, b1) T; u += w Q n, M += w Q.  Q and the
// sums are FP64: with DecorrTran for delta/beta = d, T^T T = diag(2, 2 d^2)
// has condition number d^2 (2500 at d = 50), which FP32 sums would lose.
__device__ __forceinline__ void pc_pair(const PcTab& P, int kind, float w, float Vd, float Vb,
                                        float nd, float nb, double u[2], double M[3]) {
  const float Dd = Vd - nd, Db = Vb - nb;
  const float d0 = P.T[0] * Dd + P.T[1] * Db, d1 = P.T[2] * Dd + P.T[3] * Db;
  const double b0 = w * qg_b(d0, P.qmp, P.c[kind][0], P.inv_sq[kind][0], P.inv_sqp[kind][0],
                             P.b0[kind][0]);
  const double b1 = w * qg_b(d1, P.qmp, P.c[kind][1], P.inv_sq[kind][1], P.inv_sqp[kind][1],
                             P.b0[kind][1]);
  const double t0 = P.T[0], t1 = P.T[1], t2 = P.T[2], t3 = P.T[3];
  const double q00 = t0 * b0 * t0 + t2 * b1 * t2;
  const double q01 = t0 * b0 * t1 + t2 * b1 * t3;
  const double q11 = t1 * b0 * t1 + t3 * b1 * t3;
  u[0] += q00 * nd + q01 * nb;
  u[1] += q01 * nd + q11 * nb;
  M[0] += q00;
  M[1] += q01;
  M[2] += q11;
}

__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// One pass of the tomographic ICD over the columns of the launch's tiles
// (one 2x2 tile colour, the frames of one parity), one wavefront per (tile,
// frame): per column the A window in LDS, theta2 = mu zo sum (k a)^2 (shared
// by both objects) and theta1 = -mu sum e k a of each error sinogram per
// voxel (PCMBIR_3D/XT_VoxUpdate.c:23-52), then the two z-parity halves of
// coupled substitutions with zero skipping (XT_VoxUpdate.c:208-228), then
// both error sinograms updated with atomics (XT_VoxUpdate.c:57-79).
__global__ __launch_bounds__(4 * kWave) void k_pc_sweep(SwGeom G, PcTab P, PcArgs A) {
  extern __shared__ __align__(16) float smem[];
  __shared__ float s_t1d[kNZPc], s_t1b[kNZPc], s_dxd[kNZPc], s_dxb[kNZPc];
  __shared__ float s_tn[4][kNZPc];  // delta t-1, delta t+1, beta t-1, beta t+1
  // up to four wavefronts, over the voxels of a column (pc_threads); the
  // wavefronts of a workgroup share its slab
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & (kWave - 1);
  const int wid = blockIdx.x;
  const int tile = A.launch_tiles[wid % A.n_launch_tiles];
  const int f = A.fparity + 2 * (wid / A.n_launch_tiles);
  if (f >= G.N_time) return;
  const int c0 = A.tile_ptr[tile], n = A.tile_ptr[tile + 1] - c0;
  const int s0 = A.frame_ptr[f], nv = A.frame_ptr[f + 1] - s0;
  if (n == 0 || nv == 0) return;
  // the slabs hold the tile's own T x T columns of both objects; the
  // neighbouring columns belong to tiles of other colours, which no
  // workgroup of this launch updates, so they are read from the objects
  const int Nz = G.N_z, T = A.tile, W = A.W;
  const int nx = T * T * Nz;
  float* sd = smem;
  float* sb = sd + nx;
  int* views = reinterpret_cast<int*>(sb + nx);
  int* cols = views + A.nvmax;
  int* wbase = cols + T * T;
  float* wav = reinterpret_cast<float*>(wbase + A.nvmax);
  const int oy = (tile / A.tiles_x) * T - 1, ox = (tile % A.tiles_x) * T - 1;
  const size_t vol = (size_t)G.N_y * G.N_x * Nz;
  float* xdf = A.xd + (size_t)f * vol;
  float* xbf = A.xb + (size_t)f * vol;
  for (int j = tid; j < nx; j += nth) {
    const int z = j % Nz, b = j / Nz;
    const int yy = oy + 1 + b / T, xx = ox + 1 + b % T;
    const bool in = yy < G.N_y && xx < G.N_x;
    const size_t o = in ? ((size_t)yy * G.N_x + xx) * Nz + z : 0;
    sd[j] = in ? xdf[o] : 0.f;
    sb[j] = in ? xbf[o] : 0.f;
  }
  for (int j = tid; j < n; j += nth) cols[j] = A.tile_cols[c0 + j];
  for (int j = tid; j < nv; j += nth) views[j] = A.frame_views[s0 + j];
  // seeded visiting order k -> (pb + k pa) mod n, as k_sweep
  const unsigned h = mix32(A.seed ^ mix32((unsigned)tile * 131071u + (unsigned)f * 524287u));
  int pa = 1, pb = 0;
  if (n > 1) {
    pa = 1 + (int)(h % (unsigned)(n - 1));
    while (gcd_i(pa, n) != 1) pa = pa % (n - 1) + 1;
    pb = (int)(mix32(h + 0x9e3779b9u) % (unsigned)n);
  }
  __syncthreads();
  const int Wd = G.zo;  // detector rows per voxel slice
  float sdx = 0.f, supd = 0.f, sskip = 0.f;
  int pos = pb;
  for (int k = 0; k < n; ++k) {
    const int c = cols[pos];
    pos += pa;
    if (pos >= n) pos -= n;
    const int cy = c / G.N_x, cx = c % G.N_x;
    const int ly = cy - oy, lx = cx - ox;
    const int centre0 = ((ly - 1) * T + (lx - 1)) * Nz;
    // the column's A window (values times the wavenumber) and e offsets
    for (int sl = tid; sl < nv; sl += nth) {
      const int first = A.afirst[(size_t)c * G.nslots + s0 + sl];
      wbase[sl] = first < 0 ? -1 : (views[sl] * G.N_r + first) * G.N_t;
    }
    for (int i = tid; i < nv * W; i += nth) {
      const int sl = i / W, jj = i % W;
      wav[i] = A.aval[((size_t)c * G.nslots + s0 + sl) * W + jj] * P.kw;
    }
    for (int z = tid; z < Nz; z += nth) {
      const size_t tc = (size_t)c * Nz + z;
      s_tn[0][z] = f > 0 ? A.xd[(size_t)(f - 1) * vol + tc] : 0.f;
      s_tn[1][z] = f + 1 < G.N_time ? A.xd[(size_t)(f + 1) * vol + tc] : 0.f;
      s_tn[2][z] = f > 0 ? A.xb[(size_t)(f - 1) * vol + tc] : 0.f;
      s_tn[3][z] = f + 1 < G.N_time ? A.xb[(size_t)(f + 1) * vol + tc] : 0.f;
    }
    __syncthreads();
    // theta2, shared by all voxels of the column and both objects (each
    // wavefront sums it)
    float t2 = 0.f;
    for (int i = lane; i < nv * W; i += kWave) {
      const float a = wbase[i / W] >= 0 ? wav[i] : 0.f;
      t2 += a * a;
    }
    const float th2 = P.mu * (float)Wd * wave_sum(t2);
    // theta1 of each voxel from both error sinograms: a lane per voxel, its
    // Wd detector rows of every (view, bin) entry (consecutive lanes read
    // consecutive rows)
    for (int z = tid; z < Nz; z += nth) {
      float a1 = 0.f, a2 = 0.f;
      // every entry loads, so that the loads of several entries are in
      // flight together: an entry with a = 0 (the column misses the detector
      // in that view, or a window bin past the footprint, which may lie past
      // the detector's last channel) reads the view-0, channel-0 row instead
      if (Wd == 1) {
#pragma unroll 16
        for (int i = 0; i < nv * W; ++i) {
          const int base = wbase[i / W];
          const float a = base >= 0 ? wav[i] : 0.f;
          const int o = (a != 0.f ? base + (i % W) * G.N_t : 0) + z;
          a1 -= A.ed[o] * a;
          a2 -= A.eb[o] * a;
        }
      } else {
        for (int i = 0; i < nv * W; ++i) {
          const int base = wbase[i / W];
          const float a = base >= 0 ? wav[i] : 0.f;
          const int o = (a != 0.f ? base + (i % W) * G.N_t : 0) + z * Wd;
          for (int r = 0; r < Wd; ++r) {
            a1 -= A.ed[o + r] * a;
            a2 -= A.eb[o + r] * a;
          }
        }
      }
      s_t1d[z] = P.mu * a1;
      s_t1b[z] = P.mu * a2;
    }
    __syncthreads();
    // the two z-parity halves (voxels of one parity are not neighbours)
    for (int par = 0; par < 2; ++par) {
      for (int z = par + 2 * tid; z < Nz; z += 2 * nth) {
        const int ctr = centre0 + z;
        const float Vd = sd[ctr], Vb = sb[ctr];
        double u[2] = {0.0, 0.0}, M[3] = {0.0, 0.0, 0.0};
        bool pos_nb = false;
        for (int j = 0; j < 27; ++j) {
          if (j == 13) continue;
          const int dz = j / 9 - 1, dy = (j / 3) % 3 - 1, dx = j % 3 - 1;
          const bool ok = z + dz >= 0 && z + dz < Nz && cy + dy >= 0 && cy + dy < G.N_y &&
                          cx + dx >= 0 && cx + dx < G.N_x;
          if (!ok) continue;
          const int by = ly + dy, bx = lx + dx;
          float nd, nb;
          if (by >= 1 && by <= T && bx >= 1 && bx <= T) {
            const int o = ctr + (dy * T + dx) * Nz + dz;
            nd = sd[o];
            nb = sb[o];
          } else {
            const size_t o = ((size_t)(cy + dy) * G.N_x + (cx + dx)) * Nz + z + dz;
            nd = xdf[o];
            nb = xbf[o];
          }
          pos_nb |= nd > 0.f || nb > 0.f;
          pc_pair(P, 0, P.filt[j], Vd, Vb, nd, nb, u, M);
        }
        for (int t = 0; t < 2; ++t) {
          const bool ok = t == 0 ? f > 0 : f + 1 < G.N_time;
          if (!ok) continue;
          const float nd = s_tn[t][z], nb = s_tn[2 + t][z];
          pos_nb |= nd > 0.f || nb > 0.f;
          if (P.filt[27] != 0.f) pc_pair(P, 1, P.filt[27], Vd, Vb, nd, nb, u, M);
        }
        const bool skip = Vd == 0.f && Vb == 0.f && A.iter > 1 && !pos_nb;
        float nd = Vd, nb = Vb;
        if (!skip) {
          // the 2x2 solve (th2 I + M) v = u + th2 V - th1 in FP64
          const double a = th2 + M[0], b = M[1], d = th2 + M[2];
          const double inv = 1.0 / (a * d - b * b);
          const double r0 = u[0] + (double)th2 * Vd - s_t1d[z];
          const double r1 = u[1] + (double)th2 * Vb - s_t1b[z];
          const double vd = inv * (d * r0 - b * r1), vb = inv * (a * r1 - b * r0);
          nd = (float)(Vd + P.alpha * (vd - Vd));
          nb = (float)(Vb + P.alpha * (vb - Vb));
        }
        sd[ctr] = nd;
        sb[ctr] = nb;
        s_dxd[z] = nd - Vd;
        s_dxb[z] = nb - Vb;
        sdx += sqrtf((nd - Vd) * (nd - Vd) + (nb - Vb) * (nb - Vb));
        supd += skip ? 0.f : 1.f;
        sskip += skip ? 1.f : 0.f;
      }
      __syncthreads();
    }
    // both error sinograms: e -= k a dx along the column, a lane per voxel
    // as for theta1
    for (int i = 0; i < nv * W; ++i) {
      const int base = wbase[i / W];
      const float a = base >= 0 ? wav[i] : 0.f;
      if (a == 0.f) continue;
      const int o = base + (i % W) * G.N_t;
      for (int z = tid; z < Nz; z += nth) {
        const float dd = s_dxd[z], db = s_dxb[z];
        for (int r = 0; r < Wd; ++r) {
          if (dd != 0.f) e_add(A.ed + o + z * Wd + r, -a * dd);
          if (db != 0.f) e_add(A.eb + o + z * Wd + r, -a * db);
        }
      }
    }
    __syncthreads();
  }
  // the tile's columns back (those inside the object)
  for (int j = tid; j < nx; j += nth) {
    const int z = j % Nz, b = j / Nz;
    const int yy = oy + 1 + b / T, xx = ox + 1 + b % T;
    if (yy < G.N_y && xx < G.N_x) {
      const size_t o = ((size_t)yy * G.N_x + xx) * Nz + z;
      xdf[o] = sd[j];
      xbf[o] = sb[j];
    }
  }
  sdx = wave_sum(sdx);
  supd = wave_sum(supd);
  sskip = wave_sum(sskip);
  if (lane == 0) {
    atomicAdd(&A.stats[0], (double)sdx);
    atomicAdd(&A.stats[1], (double)supd);
    atomicAdd(&A.stats[2], (double)sskip);
  }
}

// FP64 constants of the tomographic cost.
struct PcDbl {
  double T[4], c[2][2], sq[2][2], sqp[2][2], p, q, filt[28], mu;
};

__device__ __forceinline__ double pc_rho(double d, double c, double sq, double sqp, double p,
                                         double q) {
  const double a = fabs(d);
  return (pow(a, q) / sq) / (c + pow(a, q - p) / sqp);
}

// Prior term of the tomographic cost: over each voxel's 13 forward spatial
// pairs and its (t, t+1) pair, filter weight times rho_phase(T0 D) +
// rho_mag(T1 D), D the pair's (delta, beta) difference
// (PCMBIR_3D/XT_ICD_update.c:103-270); objects column-major [t][col][z].
__global__ __launch_bounds__(256) void k_pc_prior_cost(int Nt, int Ny, int Nx, int Nz, PcDbl D,
                                                       const float* __restrict__ xd,
                                                       const float* __restrict__ xb,
                                                       double* __restrict__ partial) {
  __shared__ double red[256];
  const long long nv = (long long)Nt * Ny * Nx * Nz;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nv;
       i += (long long)gridDim.x * 256) {
    const int z = (int)(i % Nz);
    const long long cz = i / Nz;
    const int col = (int)(cz % ((long long)Ny * Nx));
    const int t = (int)(cz / ((long long)Ny * Nx));
    const int yy = col / Nx, xx = col % Nx;
    const double vd = xd[i], vb = xb[i];
    for (int k = 14; k < 28; ++k) {
      long long o2;
      int kind = 0;
      if (k < 27) {
        const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
        const int z2 = z + dz, y2 = yy + dy, x2 = xx + dx;
        if (z2 < 0 || z2 >= Nz || y2 < 0 || y2 >= Ny || x2 < 0 || x2 >= Nx) continue;
        o2 = (((long long)t * Ny + y2) * Nx + x2) * Nz + z2;
      } else {
        if (t + 1 >= Nt || D.filt[27] == 0.0) continue;
        o2 = i + (long long)Ny * Nx * Nz;
        kind = 1;
      }
      const double Dd = vd - (double)xd[o2], Db = vb - (double)xb[o2];
      acc += D.filt[k] *
             (pc_rho(D.T[0] * Dd + D.T[1] * Db, D.c[kind][0], D.sq[kind][0], D.sqp[kind][0],
                     D.p, D.q) +
              pc_rho(D.T[2] * Dd + D.T[3] * Db, D.c[kind][1], D.sq[kind][1], D.sqp[kind][1],
                     D.p, D.q));
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ADMM state of the tomographic mode, per object (ENABLE_TOMO_RECONS,
// PCMBIR_3D/XT_ICD_update.c:726-760): the projection estimate Z of the phase
// retrieval, the scaled dual V, the previous estimate P.  Start: P = z,
// Z = z (or k A x, from_object: recon_type 2), V = 0, e = Z - V - k A x;
// nkax holds -k A x.
__global__ __launch_bounds__(256) void k_pc_admm_init(long long n, const double* __restrict__ z,
                                                      const float* __restrict__ nkax,
                                                      int from_object, float* __restrict__ Z,
                                                      float* __restrict__ V,
                                                      float* __restrict__ P,
                                                      float* __restrict__ e) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const float kax = -nkax[i];
    const float zz = from_object ? kax : (float)z[i];
    P[i] = (float)z[i];
    Z[i] = zz;
    V[i] = 0.f;
    e[i] = zz - kax;
  }
}

// The phase retrieval's target a = Z - e = k A x + V
// (PCMBIR_3D/XT_ICD_update.c:1089-1093).
__global__ __launch_bounds__(256) void k_pc_admm_target(long long n, const float* __restrict__ Z,
                                                        const float* __restrict__ e,
                                                        double* __restrict__ a) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    a[i] = (double)Z[i] - (double)e[i];
}

// After the phase retrieval returned znew (PCMBIR_3D/XT_ICD_update.c:1095-1140):
// e1 = znew - a, V' = -e1, e' = e1 + V - V' (so e' = Z' - V' - k A x again
// and V' = V + k A x - Z'); sums of (V' + e')^2 = (Z' - k A x)^2, of
// (Z' - P)^2 and of V'^2 per block into partial[3 block + 0..2]; P = Z = znew.
__global__ __launch_bounds__(256) void k_pc_admm_update(long long n,
                                                        const double* __restrict__ znew,
                                                        float* __restrict__ Z,
                                                        float* __restrict__ e,
                                                        float* __restrict__ V,
                                                        float* __restrict__ P,
                                                        double* __restrict__ partial) {
  __shared__ double red[3][256];
  double pr = 0.0, du = 0.0, vs = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const double zn = znew[i];
    const double a = (double)Z[i] - (double)e[i];
    const double e1 = zn - a, vold = V[i], vnew = -e1, en = e1 + vold - vnew;
    pr += (vnew + en) * (vnew + en);
    du += (zn - (double)P[i]) * (zn - (double)P[i]);
    vs += vnew * vnew;
    V[i] = (float)vnew;
    e[i] = (float)en;
    P[i] = (float)zn;
    Z[i] = (float)zn;
  }
  red[0][threadIdx.x] = pr;
  red[1][threadIdx.x] = du;
  red[2][threadIdx.x] = vs;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Data term mu/2 sum(e_d^2 + e_b^2) (PCMBIR_3D/XT_ICD_update.c:92-102 with
// mu, not mu^2: the objective the ICD minimises).
__global__ __launch_bounds__(256) void k_pc_data_cost(long long n, double mu,
                                                      const float* __restrict__ ed,
                                                      const float* __restrict__ eb,
                                                      double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    acc += (double)ed[i] * ed[i] + (double)eb[i] * eb[i];
  red[threadIdx.x] = acc * mu * 0.5;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Voxel-selection criterion of non-homogeneous ICD: the magnitude map
// filtered with the separable 5x5 Hamming window h = 0.54 - 0.46 cos(2 pi n / 4)
// = (0.08, 0.54, 1, 0.54, 0.08) per axis, truncated at the grid's edges
// (XT_NHICD.c:126-160).
__global__ __launch_bounds__(256) void k_vsc(int N_time, int N_y, int N_x,
                                             const float* __restrict__ mag,
                                             float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)N_time * N_y * N_x) return;
  const int x = (int)(i % N_x), y = (int)((i / N_x) % N_y);
  const float* m = mag + (i - ((long long)y * N_x + x));
  const float h[5] = {0.08f, 0.54f, 1.f, 0.54f, 0.08f};
  float acc = 0.f;
  for (int p = -2; p <= 2; ++p) {
    if (y + p < 0 || y + p >= N_y) continue;
    for (int q = -2; q <= 2; ++q)
      if (x + q >= 0 && x + q < N_x) acc += h[p + 2] * h[q + 2] * m[(y + p) * N_x + x + q];
  }
  out[i] = acc;
}

// Projection of the phase-contrast simulator's phantoms (ForwardProject,
// PCMBIR_3D/XT_genSinogram.c:146-186): for view p and detector sample
// (t, r), the sums over voxels of phantom x A(r) k x L(z, t) for the
// absorption and the refractive-index phantom, A the voxel's detector
// response (the A rule, times the wavenumber k) and L the voxel line
// response along z.  One workgroup per (view, chunk of detector rows): each
// thread takes voxel columns, forms the chunk's z-weighted column values
// (lz / lw: the slices feeding each row), and adds them times A into the
// chunk's rows of both projections in LDS; the rows go out once.
constexpr int kSimL = 3;    // voxel slices feeding one detector row
constexpr int kSimTc = 16;  // detector rows per workgroup at most
__global__ __launch_bounds__(256) void k_phantom_project(
    DevGeom G, const float* __restrict__ H, const double* __restrict__ cosv,
    const double* __restrict__ sinv, const float* __restrict__ mag,
    const float* __restrict__ phase, const int* __restrict__ lz, const float* __restrict__ lw,
    int n_z, int tchunk, double kwave, float* __restrict__ out_r, float* __restrict__ out_i) {
  extern __shared__ __align__(16) float smem[];
  __shared__ float sh[kNB];
  const int p = blockIdx.x, t0 = blockIdx.y * tchunk;
  const int tc = min(tchunk, n_z - t0);
  const int tid = threadIdx.x;
  float* ar = smem;
  float* ai = smem + (size_t)G.N_r * tchunk;
  for (int i = tid; i < 2 * G.N_r * tchunk; i += 256) smem[i] = 0.f;
  for (int i = tid; i < kNB; i += 256) sh[i] = H[(size_t)p * kNB + i];
  __syncthreads();
  const double cs = cosv[p], sn = sinv[p];
  const int ncol = G.N_x * G.N_y;
  for (int c = tid; c < ncol; c += 256) {
    double vr[kSimTc], vi[kSimTc];
    bool any = false;
#pragma unroll
    for (int t = 0; t < kSimTc; ++t) {
      vr[t] = 0.0;
      vi[t] = 0.0;
      if (t < tc) {
        for (int q = 0; q < kSimL; ++q) {
          const int z = lz[(t0 + t) * kSimL + q];
          if (z < 0) continue;
          const double w = lw[(t0 + t) * kSimL + q];
          vr[t] += (double)mag[(size_t)z * ncol + c] * w;
          vi[t] += (double)phase[(size_t)z * ncol + c] * w;
        }
        any |= vr[t] != 0.0 || vi[t] != 0.0;
      }
    }
    if (!any) continue;
    amatrix_visit(G, sh, cs, sn, c / G.N_x, c % G.N_x, [&](int j, int, double v) {
      const double a = v * kwave;
#pragma unroll
      for (int t = 0; t < kSimTc; ++t) {
        if (t < tc) {
          if (vr[t] != 0.0) atomicAdd(&ar[j * tchunk + t], (float)(vr[t] * a));
          if (vi[t] != 0.0) atomicAdd(&ai[j * tchunk + t], (float)(vi[t] * a));
        }
      }
    });
  }
  __syncthreads();
  for (int i = tid; i < G.N_r * tc; i += 256) {
    const int t = i / G.N_r, r = i % G.N_r;
    const size_t o = ((size_t)p * n_z + t0 + t) * G.N_r + r;
    out_r[o] = ar[r * tchunk + t];
    out_i[o] = ai[r * tchunk + t];
  }
}

// Flat-field normalisation of the APS driver (read_data,
// reconstruct/real_data/APS_Data_Format/XT_HDFIO.c:138-178).  First the
// mean of the white (or dark) frames per detector pixel: summed in float in
// frame order, divided by the frame count in float.
__global__ __launch_bounds__(256) void k_aps_mean(const uint16_t* __restrict__ f, int nf,
                                                  long long npix, float* __restrict__ out) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  float s = 0.f;
  for (int i = 0; i < nf; ++i) s = __fadd_rn(s, (float)f[(size_t)i * npix + p]);
  out[p] = __fdiv_rn(s, (float)nf);
}

// Then per output sample (view i, row j, channel k): |proj - dark| and
// |white - dark| summed in float over its ratio x ratio block of detector
// pixels (rows outer, channels inner); the weight is the first sum, the
// projection b t^2 + t with t = log(second / first) (the quotient in
// float, the log and the beam-hardening polynomial in double, unfused, the
// result stored as float).
__global__ __launch_bounds__(256) void k_aps_normalize(const uint16_t* __restrict__ proj,
                                                       const float* __restrict__ white,
                                                       const float* __restrict__ dark,
                                                       int n_proj, int rows, int cols, int ratio,
                                                       double bh, float* __restrict__ out_p,
                                                       float* __restrict__ out_w) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n_proj * rows * cols) return;
  const int k = (int)(idx % cols), j = (int)((idx / cols) % rows);
  const long long i = idx / ((long long)rows * cols);
  const long long W = (long long)cols * ratio;
  const long long frame = (long long)rows * ratio * W;
  float w = 0.f, wd = 0.f;
  for (int m = 0; m < ratio; ++m)
    for (int n = 0; n < ratio; ++n) {
      const long long q = (long long)(j * ratio + m) * W + (long long)k * ratio + n;
      const float dk = dark[q];
      w = __fadd_rn(w, fabsf(__fsub_rn((float)proj[i * frame + q], dk)));
      wd = __fadd_rn(wd, fabsf(__fsub_rn(white[q], dk)));
    }
  const double t = (double)(float)log((double)__fdiv_rn(wd, w));
  out_p[idx] = (float)__dadd_rn(__dmul_rn(__dmul_rn(bh, t), t), t);
  out_w[idx] = w;
}

// Down-sampling of the standard-format driver (read_ProjWeightData,
// reconstruct/real_data/Standard_Data_Format/XT_HDFIO.c:90-111): per output
// sample, the mean of its ratio x ratio block of projections (float sum in
// block-row-outer order, divided in float) through the beam-hardening
// polynomial b t^2 + t (double, unfused), and the sum of the block's
// weights.
__global__ __launch_bounds__(256) void k_std_downsample(const float* __restrict__ proj,
                                                        const float* __restrict__ wts,
                                                        int n_proj, int rows, int cols,
                                                        int ratio, double bh,
                                                        float* __restrict__ out_p,
                                                        float* __restrict__ out_w) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n_proj * rows * cols) return;
  const int k = (int)(idx % cols), j = (int)((idx / cols) % rows);
  const long long i = idx / ((long long)rows * cols);
  const long long W = (long long)cols * ratio;
  const long long base = i * rows * ratio * W;
  float t = 0.f, w = 0.f;
  for (int m = 0; m < ratio; ++m)
    for (int n = 0; n < ratio; ++n) {
      const long long q = base + (long long)(j * ratio + m) * W + (long long)k * ratio + n;
      t = __fadd_rn(t, proj[q]);
      w = __fadd_rn(w, wts[q]);
    }
  const double td = (double)__fdiv_rn(t, (float)(ratio * ratio));
  out_p[idx] = (float)__dadd_rn(__dmul_rn(__dmul_rn(bh, td), td), td);
  out_w[idx] = w;
}

// theta1/theta2 of individual voxels through the sweep's column routine.
template <int kW>
__global__ __launch_bounds__(kWave) void k_thetas(SwGeom G, ColumnArgs A,
                                                  const int* __restrict__ frame_ptr,
                                                  const int* __restrict__ tzyx, int nvmax,
                                                  float* __restrict__ th1,
                                                  float* __restrict__ th2) {
  extern __shared__ __align__(16) float smem[];
  __shared__ float s_row1[kNTMax], s_row2[kNTMax];
  __shared__ float s_part1[kWave], s_part2[kWave];
  __shared__ float s_tn[2 * kNZMax];
  const SweepLds L(smem, 0, nvmax, kW);
  const int i = blockIdx.x;
  const int f = tzyx[4 * i], z = tzyx[4 * i + 1];
  const int c = tzyx[4 * i + 2] * G.N_x + tzyx[4 * i + 3];
  const int s0 = frame_ptr[f], nv = frame_ptr[f + 1] - s0;
  for (int j = threadIdx.x; j < nv; j += kWave) L.views[j] = A.frame_views[s0 + j];
  wave_sync();
  APre<false> pre;
  awin_fetch<kW, false>(G, A, c, s0, nv, pre);
  for (int q = 0; q < kTnPre; ++q) pre.tn[q] = 0.f;
  awin_store<kW, false>(G, pre, nv, L.views, L.win[0], s_tn);
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync();
  col_thetas<kW>(G, A, L.win[0], nv, s_row1, s_row2, s_part1, s_part2);
  if (threadIdx.x == 0) {
    float a = 0.f, b = 0.f;
    for (int m = 0; m < G.zo; ++m) {
      a += s_row1[z * G.zo + m];
      b += s_row2[z * G.zo + m];
    }
    th1[i] = a;
    th2[i] = b;
  }
}

// Functional substitution on independent voxels with the sweep's lane split
// (16 lanes per voxel, 4 voxels per wavefront).
__global__ __launch_bounds__(kWave) void k_subst(SwGeom G, float alpha, int n,
                                                 const float* __restrict__ V,
                                                 const float* __restrict__ th1,
                                                 const float* __restrict__ th2,
                                                 const float* __restrict__ nb,
                                                 const unsigned char* __restrict__ nbf,
                                                 const float* __restrict__ tn,
                                                 const unsigned char* __restrict__ tf,
                                                 float* __restrict__ out) {
  constexpr int lpv = 16;
  __shared__ PriorTab s_prior;
  load_prior(G, &s_prior);
  __syncthreads();
  const int lane = threadIdx.x;
  const int i = blockIdx.x * (kWave / lpv) + lane / lpv;
  const int sub = lane % lpv;
  const bool act = i < n;
  const float v = act ? V[i] : 0.f;
  float num = 0.f, den = 0.f;
  if (act) {
    for (int j = sub; j < 29; j += lpv) {
      if (j == 13) continue;
      const bool ok = j < 27 ? nbf[(size_t)i * 27 + j] != 0 : tf[(size_t)i * 2 + j - 27] != 0;
      if (!ok) continue;
      const float nbv = j < 27 ? nb[(size_t)i * 27 + j] : tn[(size_t)i * 2 + j - 27];
      const float wt = nb_weight(s_prior, j, v, nbv);
      num += wt * nbv;
      den += wt;
    }
  }
  for (int off = lpv / 2; off > 0; off >>= 1) {
    num += __shfl_xor(num, off);
    den += __shfl_xor(den, off);
  }
  if (act && sub == 0) out[i] = substitute(alpha, v, th1[i], th2[i], num, den);
}

// A^T v: one thread per (column, frame), z_overlap rows folded into z.
__global__ __launch_bounds__(256) void k_backproj(DevGeom G,
                                                  const float* __restrict__ aval,
                                                  const int* __restrict__ afirst,
                                                  const int* __restrict__ frame_ptr,
                                                  const int* __restrict__ frame_views,
                                                  const float* __restrict__ v,
                                                  float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncol = (long long)G.N_x * G.N_y;
  if (i >= ncol * G.N_time) return;
  const int f = (int)(i / ncol);
  const int c = (int)(i % ncol);
  float* o = out + ((size_t)f * ncol + c) * G.N_z;
  for (int z = 0; z < G.N_z; ++z) o[z] = 0.f;
  for (int s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
    const size_t slot = (size_t)c * G.nslots + s;
    const int r0 = afirst[slot];
    if (r0 < 0) continue;
    const int p = frame_views[s];
    for (int k = 0; k < G.W; ++k) {
      const float a = aval[slot * G.W + k];
      const int r = r0 + k;
      if (a == 0.f || r >= G.N_r) continue;
      const float* vr = v + ((size_t)p * G.N_r + r) * G.N_t;
      for (int t = 0; t < G.N_t; ++t) o[t / G.zo] += a * vr[t];
    }
  }
}

// Normalised back-projection: per voxel sum(w a (y - d)) / sum(w a (A m))
// over its sinogram entries, m the update mask's indicator (A m its
// projection, `aone`): exact for a uniform object inside the mask.  The
// input of the learned warm start.  One thread per (frame, column), voxels
// column-major [frame][column][z].
__global__ __launch_bounds__(256) void k_bpnorm(DevGeom G, const float* __restrict__ aval,
                                                const int* __restrict__ afirst,
                                                const int* __restrict__ frame_ptr,
                                                const int* __restrict__ frame_views,
                                                const float* __restrict__ y,
                                                const float* __restrict__ w,
                                                const float* __restrict__ d,
                                                const float* __restrict__ aone,
                                                float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncol = (long long)G.N_x * G.N_y;
  if (i >= ncol * G.N_time) return;
  const int f = (int)(i / ncol);
  const int c = (int)(i % ncol);
  for (int z = 0; z < G.N_z; ++z) {
    double num = 0.0, den = 0.0;
    for (int s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
      const size_t slot = (size_t)c * G.nslots + s;
      const int r0 = afirst[slot];
      if (r0 < 0) continue;
      const int p = frame_views[s];
      for (int k = 0; k < G.W; ++k) {
        const float a = aval[slot * G.W + k];
        const int r = r0 + k;
        if (a == 0.f || r >= G.N_r) continue;
        const size_t row = ((size_t)p * G.N_r + r) * G.N_t;
        double sn = 0.0, sd = 0.0;
        for (int t = z * G.zo; t < (z + 1) * G.zo; ++t) {
          const double wt = w[row + t];
          sn += wt * ((double)y[row + t] - (d ? (double)d[(size_t)r * G.N_t + t] : 0.0));
          sd += wt * aone[row + t];
        }
        num += (double)a * sn;
        den += (double)a * sd;
      }
    }
    out[(size_t)i * G.N_z + z] = den > 0.0 ? (float)(num / den) : 0.f;
  }
}

// Ram-Lak filter along the detector channels of every (view, row):
// out[p][r][t] = sum_r' h(r - r') (v[p][r'][t] - d[r'][t]), h(0) = 1/4,
// h(n) = -1/(pi n)^2 for odd n, 0 for even n != 0 (the band-limited ramp
// sampled at the channel spacing).  One thread per output.
__global__ __launch_bounds__(256) void k_ramp(DevGeom G, const float* __restrict__ v,
                                              const float* __restrict__ d,
                                              float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)G.N_p * G.N_r * G.N_t;
  if (i >= n) return;
  const int t = (int)(i % G.N_t);
  const int r = (int)((i / G.N_t) % G.N_r);
  const long long p = i / ((long long)G.N_t * G.N_r);
  const float* vp = v + (size_t)p * G.N_r * G.N_t;
  double s = 0.0;
  for (int q = 0; q < G.N_r; ++q) {
    const int k = r - q;
    double h;
    if (k == 0)
      h = 0.25;
    else if (k & 1)
      h = -1.0 / (M_PI * M_PI * (double)k * (double)k);
    else
      continue;
    s += h * ((double)vp[(size_t)q * G.N_t + t] - (d ? (double)d[(size_t)q * G.N_t + t] : 0.0));
  }
  out[i] = (float)s;
}

// per-block sums of x over masked voxels (column-major object), FP64
__global__ void k_masked_sum(int Nt, int ncol, int Nz, const uint8_t* __restrict__ mask,
                             const float* __restrict__ x, double* __restrict__ partial) {
  __shared__ double red[256];
  const long long n = (long long)Nt * ncol * Nz;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    if (mask[(i / Nz) % ncol]) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// graph-safe clears and counters (no memset nodes in captured graphs)
__global__ void k_zero_f64(double* p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0) p[0] = v;
}

__global__ void k_scale(float* x, long long n, float s) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= s;
}

// Ring offset (one thread per detector element (r, t)): over the views, the
// generalized-Huber weight lambda of every entry -- w where |e sqrt(w)| < T,
// T delta sqrt(w) / |e| otherwise -- summed, and the numerator of the offset
// step: sum lambda e with lambda e = T delta sqrt(w) sign(e) outside the
// selection, or (rhs3) the right-hand side sum lambda (e + d).  FP64 sums.
__global__ void k_offset_sums(int N_p, int nrt, double T, double delta,
                              const float* __restrict__ e, const float* __restrict__ w,
                              const float* __restrict__ d, int rhs3,
                              double* __restrict__ num, double* __restrict__ lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrt) return;
  const double di = d ? (double)d[i] : 0.0;
  double a = 0.0, b = 0.0;
  for (int p = 0; p < N_p; ++p) {
    const size_t q = (size_t)p * nrt + i;
    const double ee = e[q], ww = w[q];
    const double tw = T * delta * sqrt(ww);
    if (fabs(ee * sqrt(ww)) < T) {
      b += ww;
      a += (rhs3 ? ee + di : ee) * ww;
    } else {
      const double l = tw / fabs(ee);
      b += l;
      a += rhs3 ? (ee + di) * l : tw * (double)((ee > 0) - (ee < 0));
    }
  }
  num[i] = a;
  lam[i] = b;
}

// e[p][r][t] -= step[r][t] (FP64 step, the result rounded once)
__global__ void k_offset_apply(long long n, int nrt, const double* __restrict__ step,
                               float* __restrict__ e) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) e[q] = (float)((double)e[q] - step[q % nrt]);
}

// Noise variance: per-block sums of e^2 W where |e sqrt(w)| < T, else
// |e| delta T sqrt(W var), with W = w var the measured weight (FP64).
__global__ void k_var_sums(long long n, double T, double delta, double var,
                           const float* __restrict__ e, const float* __restrict__ w,
                           double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n;
       q += (long long)gridDim.x * 256) {
    const double ee = e[q], wc = w[q];
    const double W = (float)(wc * var);
    s += fabs(ee * sqrt(wc)) < T ? ee * ee * W : fabs(ee) * delta * T * sqrt(W * var);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// w := (w var_old) / var_new, each product rounded to FP32 as the stored
// weights are
__global__ void k_var_apply(long long n, double var_old, double var_new, float* __restrict__ w) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) w[q] = (float)((double)(float)((double)w[q] * var_old) / var_new);
}

// the update mask's indicator, column-major [frame][column][z]
__global__ void k_mask_ones(int Nt, int ncol, int Nz, const uint8_t* __restrict__ mask,
                            float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)Nt * ncol * Nz) return;
  out[i] = mask[(i / Nz) % ncol] ? 1.f : 0.f;
}

// [t][z][y][x] <-> [t][y][x][z]
// Fine-grid sample i of a 2x up-sampling along an axis of n coarse samples
// (XT_ICD_update.c:358-390, 448-496): the ends copy, interior samples weigh
// the two nearest coarse samples 3/4 and 1/4.
__device__ __forceinline__ void up_axis(int i, int n, int& a0, double& w0, int& a1, double& w1) {
  const int k = i >> 1;
  if (i == 0 || i == 2 * n - 1) {
    a0 = a1 = i == 0 ? 0 : n - 1;
    w0 = 1.0;
    w1 = 0.0;
  } else if (i & 1) {
    a0 = k;
    w0 = 0.75;
    a1 = k + 1;
    w1 = 0.25;
  } else {
    a0 = k - 1;
    w0 = 0.25;
    a1 = k;
    w1 = 0.75;
  }
}

// One stage's object up-sampled into the next stage's (column-major
// [N_time][N_y * N_x][N_z] both): 2x in x and y, and in z when zf = 2; FP64
// weights, one rounding.  mask (may be null): zero outside the fine stage's
// update mask.
__global__ __launch_bounds__(256) void k_upsample(int Nt, int Nyc, int Nxc, int Nzc, int zf,
                                                  const float* __restrict__ xc,
                                                  const uint8_t* __restrict__ mask,
                                                  float* __restrict__ xf) {
  const int Nyf = 2 * Nyc, Nxf = 2 * Nxc, Nzf = zf * Nzc;
  const long long ncf = (long long)Nyf * Nxf, ncc = (long long)Nyc * Nxc;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)Nt * ncf * Nzf) return;
  const int z = (int)(i % Nzf);
  const long long c = (i / Nzf) % ncf;
  const int f = (int)(i / (ncf * Nzf));
  if (mask && !mask[c]) {
    xf[i] = 0.f;
    return;
  }
  const int y = (int)(c / Nxf), x = (int)(c % Nxf);
  int ya[2], xa[2], za[2];
  double yw[2], xw[2], zw[2];
  up_axis(y, Nyc, ya[0], yw[0], ya[1], yw[1]);
  up_axis(x, Nxc, xa[0], xw[0], xa[1], xw[1]);
  if (zf == 2) {
    up_axis(z, Nzc, za[0], zw[0], za[1], zw[1]);
  } else {
    za[0] = za[1] = z;
    zw[0] = 1.0;
    zw[1] = 0.0;
  }
  const float* src = xc + (size_t)f * ncc * Nzc;
  double v = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int d = 0; d < 2; ++d) {
        const double w = yw[a] * xw[b] * zw[d];
        if (w != 0.0) v += w * src[((size_t)ya[a] * Nxc + xa[b]) * Nzc + za[d]];
      }
  xf[i] = (float)v;
}

__global__ __launch_bounds__(256) void k_to_colmajor(int Nt, int Nz, int ncol,
                                                     const float* __restrict__ in,
                                                     float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = (long long)Nt * Nz * ncol;
  if (i >= n) return;
  const int c = (int)(i % ncol);
  const int z = (int)((i / ncol) % Nz);
  const int f = (int)(i / ((long long)ncol * Nz));
  out[((size_t)f * ncol + c) * Nz + z] = in[i];
}

__global__ __launch_bounds__(256) void k_from_colmajor(int Nt, int Nz, int ncol,
                                                       const float* __restrict__ in,
                                                       float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = (long long)Nt * Nz * ncol;
  if (i >= n) return;
  const int c = (int)(i % ncol);
  const int z = (int)((i / ncol) % Nz);
  const int f = (int)(i / ((long long)ncol * Nz));
  out[i] = in[((size_t)f * ncol + c) * Nz + z];
}

DevGeom make_dev_geom(const mbir_hip_geom& g, int nslots) {
  DevGeom G;
  memset(&G, 0, sizeof(G));
  G.N_r = g.N_r;
  G.N_t = g.N_t;
  G.N_p = g.N_p;
  G.N_x = g.N_x;
  G.N_y = g.N_y;
  G.N_z = g.N_z;
  G.N_time = g.N_time;
  G.W = 2 * g.mult_xy + 1;
  G.zo = g.mult_z;
  G.nslots = nslots;
  G.delta_r = g.delta_r;
  G.R0 = g.R0;
  G.RMax = g.RMax;
  G.x0 = g.x0;
  G.y0 = g.y0;
  G.delta_xy = g.delta_xy;
  G.OffsetR = g.OffsetR;
  G.delta_t = (float)g.delta_t;
  G.q = (float)g.q;
  G.p_s = (float)g.p_s;
  G.p_t = (float)g.p_t;
  G.c_s = (float)g.c_s;
  G.c_t = (float)g.c_t;
  G.inv_sq_s = (float)(1.0 / g.sig_s_q);
  G.inv_sqp_s = (float)(1.0 / g.sig_s_qp);
  G.inv_sq_t = (float)(1.0 / g.sig_t_q);
  G.inv_sqp_t = (float)(1.0 / g.sig_t_qp);
  G.b0_s = (float)(g.q / (g.sig_s_q * g.c_s));
  G.b0_t = (float)(g.q / (g.sig_t_q * g.c_t));
  for (int k = 0; k < 27; ++k) {
    G.filt[k] = (float)g.filt[k];
    G.dfilt[k] = g.filt[k];
  }
  G.filt_t = (float)g.filt_t;
  G.alpha = (float)g.alpha;
  G.huber_T = (float)g.huber_T;
  G.huber_delta = (float)g.huber_delta;
  G.dq = g.q;
  G.dp_s = g.p_s;
  G.dp_t = g.p_t;
  G.dc_s = g.c_s;
  G.dc_t = g.c_t;
  G.dsq_s = g.sig_s_q;
  G.dsqp_s = g.sig_s_qp;
  G.dsq_t = g.sig_t_q;
  G.dsqp_t = g.sig_t_qp;
  G.dfilt_t = g.filt_t;
  G.dhuber_T = g.huber_T;
  G.dhuber_delta = g.huber_delta;
  return G;
}

SwGeom make_sw_geom(const mbir_hip_geom& g, int nslots) {
  SwGeom S;
  memset(&S, 0, sizeof(S));
  S.N_r = g.N_r;
  S.N_t = g.N_t;
  S.N_x = g.N_x;
  S.N_y = g.N_y;
  S.N_z = g.N_z;
  S.N_time = g.N_time;
  S.zo = g.mult_z;
  S.nslots = nslots;
  S.delta_t = (float)g.delta_t;
  S.huber_T = (float)g.huber_T;
  S.huber_delta = (float)g.huber_delta;
  const double p[2] = {g.p_s, g.p_t}, c[2] = {g.c_s, g.c_t};
  const double sq[2] = {g.sig_s_q, g.sig_t_q}, sqp[2] = {g.sig_s_qp, g.sig_t_qp};
  for (int t = 0; t < 2; ++t) {
    S.prior.qmp[t] = (float)(g.q - p[t]);
    S.prior.c[t] = (float)c[t];
    S.prior.inv_sq[t] = (float)(1.0 / sq[t]);
    S.prior.inv_sqp[t] = (float)(1.0 / sqp[t]);
    S.prior.b0[t] = (float)(g.q / (sq[t] * c[t]));
  }
  for (int k = 0; k < 27; ++k) S.prior.filt[k] = (float)g.filt[k];
  S.prior.filt[27] = (float)g.filt_t;
  return S;
}

template <typename T>
int dev_alloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIP_OK(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

}  // namespace

struct mbir_hip_ctx {
  mbir_hip_geom g;
  DevGeom G;
  SwGeom SG;
  hipStream_t stream = nullptr;
  int nslots = 0, ncol = 0, nvmax = 0;
  size_t nsino = 0, nvox = 0;
  float* d_H = nullptr;
  double *d_cos = nullptr, *d_sin = nullptr;
  int *d_frame_ptr = nullptr, *d_frame_views = nullptr, *d_view_frame = nullptr;
  float *d_y = nullptr, *d_w = nullptr, *d_d = nullptr, *d_e = nullptr;
  float *d_x = nullptr, *d_stage = nullptr, *d_stage2 = nullptr;
  uint8_t* d_mask = nullptr;  // update mask [N_y * N_x]
  float* d_ws = nullptr;      // warm-start sinogram scratch (allocated on first use)
  std::vector<float> offset;  // ring offset d [N_r][N_t] (host copy of d_d)
  // ring-offset windows (mode 3): per sinogram entry the windows covering it
  // and their weights, [off_cover_ptr[k], off_cover_ptr[k + 1])
  int off_windows = 0;
  std::vector<int> off_cover_ptr, off_cover_win;
  std::vector<double> off_cover_val;
  double* d_rt = nullptr;     // [3][N_r N_t] FP64 sums and step of the offset update
  // captured equits (mbir_hip_icd_graph): the graph, its key and launches
  int* d_iter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long gkey[7] = {-1, -1, -1, -1, -1, -1, -1};
  int glaunches = 0;
  // non-homogeneous ICD (option "nhicd"): per (frame, column) sum of |dx| of
  // the column's last visit and its VSC-filtered ranking, [N_time][N_y N_x]
  bool nh = false;
  float *d_mag = nullptr, *d_vsc = nullptr;
  int* d_afirst = nullptr;
  float* d_aval = nullptr;
  bool afp16 = false;                 // the sweep streams FP16 A + column scales
  unsigned short* d_aval16 = nullptr;  // [ncol][nslots][W] FP16 bits
  float* d_ascale = nullptr;           // [ncol]
  bool a_built = false;
  int *d_tile_ptr = nullptr, *d_tile_cols = nullptr, *d_launch_tiles = nullptr;
  int launch_off[5] = {0, 0, 0, 0, 0};
  int tile = 8, tiles_for = -1;
  double alpha = -1.0;
  size_t lds_max = 64 * 1024;  // LDS bytes one workgroup may allocate
  // sweep placement (option "placement"): 0 tile-major grid, 1 frames bound
  // to XCDs
  int placement = 0;
  unsigned* d_xcd = nullptr;  // [items run, items off their XCD] (option "xcd_check")
  bool pair = false;          // option "pair_parities"
  float* d_snap = nullptr;    // saved object (mbir_hip_object_snapshot)
  bool snap_valid = false;
  // PCMBIR tomographic mode (mbir_hip_pc_enable): d_x holds delta, d_x2
  // beta; d_e / d_e2 their error sinograms; d_tgt staging for targets
  bool pc = false;
  float *d_x2 = nullptr, *d_e2 = nullptr;
  PcTab ptab;
  PcDbl pdbl;
  // its ADMM state (mbir_hip_pc_admm_*): Z, V, P of delta then beta, 2 nsino
  // each; d_pk -k A x of one object; d_pdbl an FP64 sinogram of staging
  float *d_pz = nullptr, *d_pv = nullptr, *d_pp = nullptr, *d_pk = nullptr;
  double* d_pdbl = nullptr;
  bool admm = false;
  double* d_stats = nullptr;
  double* d_partial = nullptr;
  size_t partial_cap = 0;
  std::vector<uint8_t> mask;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[MBIR_K_COUNT];
  size_t ev_used[MBIR_K_COUNT] = {0, 0, 0, 0, 0};
  double ms[MBIR_K_COUNT] = {0, 0, 0, 0, 0};
  int64_t launches[MBIR_K_COUNT] = {0, 0, 0, 0, 0};
};

namespace {

int ev_begin(mbir_hip_ctx* c, int k) {
  if (c->ev_used[k] == c->ev[k].size()) {
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    c->ev[k].push_back({a, b});
  }
  HIP_OK(hipEventRecord(c->ev[k][c->ev_used[k]].first, c->stream));
  return 0;
}

int ev_end(mbir_hip_ctx* c, int k) {
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(c->ev[k][c->ev_used[k]].second, c->stream));
  c->ev_used[k]++;
  c->launches[k]++;
  return 0;
}

int ev_collect(mbir_hip_ctx* c) {
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < MBIR_K_COUNT; ++k) {
    for (size_t i = 0; i < c->ev_used[k]; ++i) {
      float m = 0.f;
      HIP_OK(hipEventElapsedTime(&m, c->ev[k][i].first, c->ev[k][i].second));
      c->ms[k] += m;
    }
    c->ev_used[k] = 0;
  }
  return 0;
}

int ensure_partial(mbir_hip_ctx* c, size_t n) {
  if (n <= c->partial_cap) return 0;
  if (c->d_partial) HIP_OK(hipFree(c->d_partial));
  if (dev_alloc(&c->d_partial, n)) return -1;
  c->partial_cap = n;
  return 0;
}

// Tiles of tile x tile columns, masked columns only, grouped by 2x2 colour.
int build_tiles(mbir_hip_ctx* c) {
  if (c->tiles_for == c->tile) return 0;
  const int T = c->tile, Nx = c->g.N_x, Ny = c->g.N_y;
  const int TX = (Nx + T - 1) / T, TY = (Ny + T - 1) / T;
  std::vector<int> ptr(1, 0), cols, order[4];
  for (int ty = 0; ty < TY; ++ty)
    for (int tx = 0; tx < TX; ++tx) {
      const int id = ty * TX + tx;
      for (int yy = ty * T; yy < std::min(Ny, ty * T + T); ++yy)
        for (int xx = tx * T; xx < std::min(Nx, tx * T + T); ++xx)
          if (c->mask[(size_t)yy * Nx + xx]) cols.push_back(yy * Nx + xx);
      ptr.push_back((int)cols.size());
      if (ptr[id + 1] > ptr[id]) order[(tx % 2) + 2 * (ty % 2)].push_back(id);
    }
  std::vector<int> launch;
  for (int k = 0; k < 4; ++k) {
    c->launch_off[k] = (int)launch.size();
    launch.insert(launch.end(), order[k].begin(), order[k].end());
  }
  c->launch_off[4] = (int)launch.size();
  if (c->d_tile_ptr) HIP_OK(hipFree(c->d_tile_ptr));
  if (c->d_tile_cols) HIP_OK(hipFree(c->d_tile_cols));
  if (c->d_launch_tiles) HIP_OK(hipFree(c->d_launch_tiles));
  if (dev_alloc(&c->d_tile_ptr, ptr.size()) || dev_alloc(&c->d_tile_cols, cols.size()) ||
      dev_alloc(&c->d_launch_tiles, launch.size()))
    return -1;
  HIP_OK(hipMemcpy(c->d_tile_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!cols.empty())
    HIP_OK(hipMemcpy(c->d_tile_cols, cols.data(), cols.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  if (!launch.empty())
    HIP_OK(hipMemcpy(c->d_launch_tiles, launch.data(), launch.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  c->tiles_for = T;
  return 0;
}

size_t fwd_tchunk(const mbir_hip_geom& g, size_t acc_bytes) {
  const size_t cap = 56 * 1024;  // LDS budget per workgroup (dynamic part)
  size_t tc = cap / (acc_bytes * (size_t)g.N_r);
  if (tc > (size_t)g.N_t) tc = g.N_t;
  // at least about 1024 workgroups (four per CU; 2048 measured slower): a
  // view's rows are split further when there are few views and rows, at the
  // price of computing each column's A window once per chunk
  const size_t par = ((size_t)g.N_t * g.N_p + 1023) / 1024;
  if (tc > par) tc = par;
  if (tc < 1) tc = 1;
  return tc;
}

// mode 0: out = y - A x - d, mode 1: data cost partials, mode 2: out = A x.
// y (mode 0) replaces the scan's projections and drops the offset d, and
// xscale multiplies A (the tomographic ICD's error sinograms: y the target,
// xscale the wavenumber).
int run_fwdproj(mbir_hip_ctx* c, int mode, const float* x, float* out,
                const float* y = nullptr, double xscale = 1.0) {
  const mbir_hip_geom& g = c->g;
  // FP64 accumulation for the error sinogram and the cost, so e is the
  // correctly rounded y - A x - d however long the sweeps have run
  const bool dbl = (mode == 0 || mode == 1);
  const size_t tc = fwd_tchunk(g, dbl ? 8 : 4);
  if ((size_t)g.N_r * (dbl ? 8 : 4) > 56 * 1024)
    return fail("detector too wide for the LDS projection accumulator");
  const int nch = (int)((g.N_t + tc - 1) / tc);
  dim3 grid(g.N_p, nch);
  const size_t shm = (size_t)g.N_r * tc * (dbl ? 8 : 4);
  if (mode == 1 && ensure_partial(c, (size_t)g.N_p * nch)) return -1;
  if (ev_begin(c, mode == 1 ? MBIR_K_COST : MBIR_K_FWDPROJ)) return -1;
  if (dbl)
    hipLaunchKernelGGL(k_fwdproj<double>, grid, dim3(256), shm, c->stream, c->G, mode,
                       (int)tc, c->d_H, c->d_cos, c->d_sin, c->d_view_frame, x,
                       y ? y : c->d_y, c->d_w, y ? nullptr : c->d_d, out, c->d_partial,
                       xscale);
  else
    hipLaunchKernelGGL(k_fwdproj<float>, grid, dim3(256), shm, c->stream, c->G, mode,
                       (int)tc, c->d_H, c->d_cos, c->d_sin, c->d_view_frame, x,
                       y ? y : c->d_y, c->d_w, y ? nullptr : c->d_d, out, c->d_partial,
                       xscale);
  return ev_end(c, mode == 1 ? MBIR_K_COST : MBIR_K_FWDPROJ);
}

int sum_partials(mbir_hip_ctx* c, size_t n, double* out) {
  std::vector<double> h(n);
  HIP_OK(hipMemcpyAsync(h.data(), c->d_partial, n * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (double v : h) s += v;
  *out = s;
  return 0;
}

int check_geom(const mbir_hip_geom* g) {
  if (!g) return fail("null geometry");
  if (g->N_r <= 0 || g->N_t <= 0 || g->N_p <= 0 || g->N_x <= 0 || g->N_y <= 0 ||
      g->N_z <= 0 || g->N_time <= 0 || g->mult_xy <= 0 || g->mult_z <= 0)
    return fail("non-positive geometry size");
  if (g->N_z > kNZPc) return fail("N_z exceeds 256 voxels per column");
  if (g->N_t > kNTMax) return fail("N_t exceeds the sweep kernel's LDS rows (256)");
  if (2 * g->mult_xy + 1 > kWMax) return fail("mult_xy exceeds 16");
  if (g->q != 2.0) return fail("q must be 2 (the qGGMRF of the sweep kernels)");
  if ((long long)g->N_p * g->N_r * g->N_t >= (1LL << 31))
    return fail("sinogram too large for 32-bit offsets");
  if ((long long)g->N_z * g->mult_z > g->N_t) return fail("N_z*mult_z exceeds N_t");
  return 0;
}

}  // namespace

extern "C" {

const char* mbir_hip_last_error(void) { return g_err.c_str(); }

int mbir_hip_create(mbir_hip_ctx** out, const mbir_hip_geom* g, const float* H,
                    const double* cosv, const double* sinv, const int32_t* frame_ptr,
                    const int32_t* frame_views, const float* y, const float* w,
                    const float* d, const uint8_t* mask) {
  if (!out) return fail("null output pointer");
  *out = nullptr;
  if (check_geom(g)) return -1;
  if (!H || !cosv || !sinv || !frame_ptr || !frame_views || !y || !w || !mask)
    return fail("null input array");
  auto* c = new mbir_hip_ctx();
  c->g = *g;
  c->nslots = frame_ptr[g->N_time];
  c->ncol = g->N_x * g->N_y;
  c->nsino = (size_t)g->N_p * g->N_r * g->N_t;
  c->nvox = (size_t)g->N_time * g->N_z * c->ncol;
  c->G = make_dev_geom(*g, c->nslots);
  c->SG = make_sw_geom(*g, c->nslots);
  c->mask.assign(mask, mask + c->ncol);
  for (int f = 0; f < g->N_time; ++f)
    if (frame_ptr[f + 1] < frame_ptr[f]) {
      delete c;
      return fail("frame_ptr not monotone");
    }
  for (int f = 0; f < g->N_time; ++f)
    c->nvmax = std::max(c->nvmax, (int)(frame_ptr[f + 1] - frame_ptr[f]));
  std::vector<int> view_frame(g->N_p, -1);
  for (int f = 0; f < g->N_time; ++f)
    for (int s = frame_ptr[f]; s < frame_ptr[f + 1]; ++s) {
      if (frame_views[s] < 0 || frame_views[s] >= g->N_p) {
        delete c;
        return fail("frame_views entry out of range");
      }
      view_frame[frame_views[s]] = f;
    }
  int rc = 0;
  int dev = 0, lds = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      lds > 0)
    c->lds_max = (size_t)lds;
  if (hipStreamCreate(&c->stream) != hipSuccess) rc = fail("hipStreamCreate failed");
  rc = rc || dev_alloc(&c->d_H, (size_t)g->N_p * kNB) || dev_alloc(&c->d_cos, g->N_p) ||
       dev_alloc(&c->d_sin, g->N_p) || dev_alloc(&c->d_frame_ptr, g->N_time + 1) ||
       dev_alloc(&c->d_frame_views, c->nslots) || dev_alloc(&c->d_view_frame, g->N_p) ||
       dev_alloc(&c->d_y, c->nsino) || dev_alloc(&c->d_w, c->nsino) ||
       dev_alloc(&c->d_e, c->nsino) || dev_alloc(&c->d_x, c->nvox) ||
       dev_alloc(&c->d_stage, std::max(c->nvox, c->nsino)) ||
       dev_alloc(&c->d_stage2, std::max(c->nvox, c->nsino)) ||
       dev_alloc(&c->d_stats, 4) || dev_alloc(&c->d_mask, (size_t)c->ncol);
  if (!rc && d) rc = dev_alloc(&c->d_d, (size_t)g->N_r * g->N_t);
  if (rc) {
    std::string keep = g_err;
    mbir_hip_destroy(c);
    return fail(keep);
  }
#define UP(dst, src, n)                                                           \
  if (hipMemcpy(dst, src, (n) * sizeof(*(src)), hipMemcpyHostToDevice) != hipSuccess) { \
    mbir_hip_destroy(c);                                                          \
    return fail("upload of " #src " failed");                                     \
  }
  UP(c->d_H, H, (size_t)g->N_p * kNB);
  UP(c->d_cos, cosv, g->N_p);
  UP(c->d_sin, sinv, g->N_p);
  UP(c->d_frame_ptr, frame_ptr, g->N_time + 1);
  if (c->nslots) UP(c->d_frame_views, frame_views, c->nslots);
  UP(c->d_view_frame, view_frame.data(), g->N_p);
  UP(c->d_y, y, c->nsino);
  UP(c->d_w, w, c->nsino);
  UP(c->d_mask, mask, (size_t)c->ncol);
  if (d) UP(c->d_d, d, (size_t)g->N_r * g->N_t);
  c->offset.assign((size_t)g->N_r * g->N_t, 0.f);
  if (d) c->offset.assign(d, d + (size_t)g->N_r * g->N_t);
#undef UP
  if (hipMemset(c->d_x, 0, c->nvox * sizeof(float)) != hipSuccess ||
      hipMemset(c->d_e, 0, c->nsino * sizeof(float)) != hipSuccess) {
    mbir_hip_destroy(c);
    return fail("hipMemset failed");
  }
  *out = c;
  return 0;
}

void mbir_hip_destroy(mbir_hip_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->d_H, c->d_cos, c->d_sin, c->d_frame_ptr, c->d_frame_views,
                  c->d_view_frame, c->d_y, c->d_w, c->d_d, c->d_e, c->d_x,
                  c->d_stage, c->d_stage2, c->d_afirst, c->d_aval, c->d_tile_ptr,
                  c->d_tile_cols, c->d_launch_tiles, c->d_stats, c->d_partial, c->d_mask,
                  c->d_aval16, c->d_ascale, c->d_ws, c->d_rt, c->d_mag, c->d_vsc,
                  c->d_xcd, c->d_snap, c->d_x2, c->d_e2, c->d_pz, c->d_pv, c->d_pp,
                  c->d_pk, c->d_pdbl};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->d_iter) (void)hipFree(c->d_iter);
  for (int k = 0; k < MBIR_K_COUNT; ++k)
    for (auto& pr : c->ev[k]) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mbir_hip_set_option(mbir_hip_ctx* c, const char* key, double value) {
  if (!c || !key) return fail("null argument");
  const std::string k(key);
  if (k == "tile") {
    // tiles of 2x2 columns make the sweep diverge: that many concurrent
    // columns (a quarter of the tiles, 4 columns each), each reading an
    // error sinogram that misses the others' latest updates, overshoot even
    // at alpha 1 (the oracle's asynchronous model shows it with 1-3 columns
    // of staleness, and convergence at tile 4), so they are refused
    if (value < 4 || value > 1024) return fail("tile must be in [4, 1024]");
    c->tile = (int)value;
  } else if (k == "alpha") {
    c->alpha = value;
  } else if (k == "afp16") {
    if (value != 0 && value != 1) return fail("afp16 must be 0 or 1");
    if (c->afp16 != (value == 1)) c->a_built = false;  // A must be built again
    c->afp16 = value == 1;
  } else if (k == "nhicd") {
    if (value != 0 && value != 1) return fail("nhicd must be 0 or 1");
    if (value == 1 && !c->d_mag) {
      const size_t nm = (size_t)c->ncol * c->g.N_time;
      if (dev_alloc(&c->d_mag, nm) || dev_alloc(&c->d_vsc, nm)) return -1;
      HIP_OK(hipMemset(c->d_mag, 0, nm * sizeof(float)));
    }
    c->nh = value == 1;
  } else if (k == "placement") {
    if (value != 0 && value != 1) return fail("placement must be 0 or 1");
    c->placement = (int)value;
  } else if (k == "pair_parities") {
    // four launches per equit, (colour k, even frames) with (colour k + 1,
    // odd frames) in each (placement 0 only)
    if (value != 0 && value != 1) return fail("pair_parities must be 0 or 1");
    c->pair = value == 1;
  } else if (k == "xcd_check") {
    // count the sweep's work items and those that ran off their XCD
    if (value != 0 && value != 1) return fail("xcd_check must be 0 or 1");
    if (value == 1 && !c->d_xcd) {
      if (dev_alloc(&c->d_xcd, 2)) return -1;
      HIP_OK(hipMemset(c->d_xcd, 0, 2 * sizeof(unsigned)));
    }
    if (value == 0 && c->d_xcd) {
      HIP_OK(hipStreamSynchronize(c->stream));
      HIP_OK(hipFree(c->d_xcd));
      c->d_xcd = nullptr;
    }
  } else {
    return fail("unknown option " + k);
  }
  return 0;
}

int mbir_hip_set_object(mbir_hip_ctx* c, const float* x) {
  if (!c || !x) return fail("null argument");
  HIP_OK(hipMemcpyAsync(c->d_stage, x, c->nvox * sizeof(float), hipMemcpyHostToDevice,
                        c->stream));
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  hipLaunchKernelGGL(k_to_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_stage, c->d_x);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_upsample_object(mbir_hip_ctx* fine, mbir_hip_ctx* coarse, int32_t masked) {
  if (!fine || !coarse) return fail("null context");
  const mbir_hip_geom& F = fine->g;
  const mbir_hip_geom& Cg = coarse->g;
  const int zf = F.N_z == Cg.N_z ? 1 : (F.N_z == 2 * Cg.N_z ? 2 : 0);
  if (F.N_time != Cg.N_time || F.N_y != 2 * Cg.N_y || F.N_x != 2 * Cg.N_x || zf == 0)
    return fail("the fine stage must double x and y (and keep or double z) of the coarse stage");
  HIP_OK(hipStreamSynchronize(coarse->stream));
  const long long n = (long long)fine->nvox;
  hipLaunchKernelGGL(k_upsample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, fine->stream,
                     F.N_time, Cg.N_y, Cg.N_x, Cg.N_z, zf, coarse->d_x,
                     masked ? fine->d_mask : nullptr, fine->d_x);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(fine->stream));
  return 0;
}

int mbir_hip_get_object(mbir_hip_ctx* c, float* x) {
  if (!c || !x) return fail("null argument");
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  hipLaunchKernelGGL(k_from_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_x, c->d_stage);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(x, c->d_stage, c->nvox * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_object_snapshot(mbir_hip_ctx* c, int32_t restore) {
  if (!c) return fail("null context");
  if (restore != 0 && restore != 1) return fail("restore must be 0 or 1");
  if (restore) {
    if (!c->snap_valid) return fail("no saved object to restore");
    HIP_OK(hipMemcpyAsync(c->d_x, c->d_snap, c->nvox * sizeof(float), hipMemcpyDeviceToDevice,
                          c->stream));
    return 0;
  }
  if (!c->d_snap && dev_alloc(&c->d_snap, c->nvox)) return -1;
  HIP_OK(hipMemcpyAsync(c->d_snap, c->d_x, c->nvox * sizeof(float), hipMemcpyDeviceToDevice,
                        c->stream));
  c->snap_valid = true;
  return 0;
}

int mbir_hip_build_amatrix(mbir_hip_ctx* c) {
  if (!c) return fail("null context");
  const size_t n = (size_t)c->ncol * c->nslots;
  if (!c->d_afirst) {
    if (dev_alloc(&c->d_afirst, n) || dev_alloc(&c->d_aval, n * c->G.W)) return -1;
  }
  if (n == 0) return 0;
  if (ev_begin(c, MBIR_K_AMATRIX)) return -1;
  hipLaunchKernelGGL(k_amatrix, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->G, c->d_H, c->d_cos, c->d_sin, c->d_frame_views, c->d_afirst,
                     c->d_aval);
  if (c->afp16) {
    if (!c->d_aval16 &&
        (dev_alloc(&c->d_aval16, n * c->G.W) || dev_alloc(&c->d_ascale, (size_t)c->ncol)))
      return -1;
    hipLaunchKernelGGL(k_aquant, dim3((unsigned)((c->ncol + 3) / 4)), dim3(256), 0, c->stream,
                       c->ncol, c->nslots * c->G.W, c->d_aval, c->d_aval16, c->d_ascale);
  }
  if (ev_end(c, MBIR_K_AMATRIX)) return -1;
  c->a_built = true;
  return 0;
}

int mbir_hip_get_amatrix(mbir_hip_ctx* c, int32_t* first, float* val) {
  if (!c || !first || !val) return fail("null argument");
  if (!c->a_built) return fail("A not built");
  const size_t n = (size_t)c->ncol * c->nslots;
  HIP_OK(hipMemcpyAsync(first, c->d_afirst, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(val, c->d_aval, n * c->G.W * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_init_error(mbir_hip_ctx* c) {
  if (!c) return fail("null context");
  return run_fwdproj(c, 0, c->d_x, c->d_e);
}

int mbir_hip_get_error(mbir_hip_ctx* c, float* e) {
  if (!c || !e) return fail("null argument");
  HIP_OK(hipMemcpyAsync(e, c->d_e, c->nsino * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_set_error(mbir_hip_ctx* c, const float* e) {
  if (!c || !e) return fail("null argument");
  HIP_OK(hipMemcpyAsync(c->d_e, e, c->nsino * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_forward_project(mbir_hip_ctx* c, const float* x, float* ax) {
  if (!c || !x || !ax) return fail("null argument");
  HIP_OK(hipMemcpyAsync(c->d_stage, x, c->nvox * sizeof(float), hipMemcpyHostToDevice,
                        c->stream));
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  hipLaunchKernelGGL(k_to_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_stage, c->d_stage2);
  HIP_OK(hipGetLastError());
  if (run_fwdproj(c, 2, c->d_stage2, c->d_stage)) return -1;
  HIP_OK(hipMemcpyAsync(ax, c->d_stage, c->nsino * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_backproject(mbir_hip_ctx* c, const float* v, float* bt) {
  if (!c || !v || !bt) return fail("null argument");
  if (!c->a_built) return fail("A not built");
  HIP_OK(hipMemcpyAsync(c->d_stage, v, c->nsino * sizeof(float), hipMemcpyHostToDevice,
                        c->stream));
  const long long n = (long long)c->ncol * c->g.N_time;
  if (ev_begin(c, MBIR_K_BACKPROJ)) return -1;
  hipLaunchKernelGGL(k_backproj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->G, c->d_aval, c->d_afirst, c->d_frame_ptr, c->d_frame_views,
                     c->d_stage, c->d_stage2);
  if (ev_end(c, MBIR_K_BACKPROJ)) return -1;
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  hipLaunchKernelGGL(k_from_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_stage2, c->d_stage);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(bt, c->d_stage, c->nvox * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_bp_normalized(mbir_hip_ctx* c, float* out) {
  if (!c || !out) return fail("null argument");
  if (!c->a_built) return fail("A not built");
  const long long n = (long long)c->ncol * c->g.N_time;
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  // A m into d_stage, then the ratio into d_stage2
  hipLaunchKernelGGL(k_mask_ones, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time, c->ncol,
                     c->g.N_z, c->d_mask, c->d_stage2);
  if (run_fwdproj(c, 2, c->d_stage2, c->d_stage)) return -1;
  hipLaunchKernelGGL(k_bpnorm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->G,
                     c->d_aval, c->d_afirst, c->d_frame_ptr, c->d_frame_views, c->d_y, c->d_w,
                     c->d_d, c->d_stage, c->d_stage2);
  hipLaunchKernelGGL(k_from_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_stage2, c->d_stage);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, c->d_stage, c->nvox * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_fbp(mbir_hip_ctx* c, float* out) {
  if (!c || !out) return fail("null argument");
  if (!c->a_built) return fail("A not built");
  if (!c->d_ws && dev_alloc(&c->d_ws, c->nsino)) return -1;
  const long long ncf = (long long)c->ncol * c->g.N_time;
  const unsigned bo = (unsigned)((c->nvox + 255) / 256);
  const unsigned bs = (unsigned)((c->nsino + 255) / 256);
  const unsigned bc = (unsigned)((ncf + 255) / 256);
  // K = mean over the mask of A^T ramp(A m): what a uniform object of value 1
  // inside the mask reconstructs to
  hipLaunchKernelGGL(k_mask_ones, dim3(bo), dim3(256), 0, c->stream, c->g.N_time, c->ncol,
                     c->g.N_z, c->d_mask, c->d_stage2);
  if (run_fwdproj(c, 2, c->d_stage2, c->d_stage)) return -1;
  hipLaunchKernelGGL(k_ramp, dim3(bs), dim3(256), 0, c->stream, c->G, c->d_stage, nullptr, c->d_ws);
  hipLaunchKernelGGL(k_backproj, dim3(bc), dim3(256), 0, c->stream, c->G, c->d_aval, c->d_afirst,
                     c->d_frame_ptr, c->d_frame_views, c->d_ws, c->d_stage2);
  const unsigned nb = 1024;
  if (ensure_partial(c, nb)) return -1;
  hipLaunchKernelGGL(k_masked_sum, dim3(nb), dim3(256), 0, c->stream, c->g.N_time, c->ncol,
                     c->g.N_z, c->d_mask, c->d_stage2, c->d_partial);
  double sum = 0.0;
  if (sum_partials(c, nb, &sum)) return -1;
  size_t nmask = 0;
  for (uint8_t m : c->mask) nmask += m ? 1 : 0;
  const double K = sum / ((double)nmask * c->g.N_z * c->g.N_time);
  if (!(K > 0.0)) return fail("filtered back-projection of the mask is not positive");
  // A^T ramp(y - d) / K
  hipLaunchKernelGGL(k_ramp, dim3(bs), dim3(256), 0, c->stream, c->G, c->d_y, c->d_d, c->d_ws);
  hipLaunchKernelGGL(k_backproj, dim3(bc), dim3(256), 0, c->stream, c->G, c->d_aval, c->d_afirst,
                     c->d_frame_ptr, c->d_frame_views, c->d_ws, c->d_stage2);
  hipLaunchKernelGGL(k_scale, dim3(bo), dim3(256), 0, c->stream, c->d_stage2, (long long)c->nvox,
                     (float)(1.0 / K));
  hipLaunchKernelGGL(k_from_colmajor, dim3(bo), dim3(256), 0, c->stream, c->g.N_time, c->g.N_z,
                     c->ncol, c->d_stage2, c->d_stage);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, c->d_stage, c->nvox * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

namespace {

// The MBIR sweep's per-column limits (its LDS column, window and register
// staging; the tomographic sweep has its own): 0 when the stage fits.
int mbir_sweep_fits(const mbir_hip_ctx* c) {
  if (c->g.N_z > kNZMax) return fail("the MBIR sweep takes at most 128 voxels per column");
  if (c->nvmax > kMaxViews || c->nvmax * (2 * c->g.mult_xy + 1) > kMaxAV)
    return fail("too many views per frame for the MBIR sweep's LDS window (512 views, "
                "1024 A values)");
  return 0;
}

// The 8 launches per equit (2x2 tile colours x 2 frame parities) of n equits.
// iter_dev null: iterations first_iter .. first_iter + n - 1, each launch
// timed; otherwise the iteration is *iter_dev + it (graph capture: no event
// records).  Returns the number of launches.
int sweep_launches(mbir_hip_ctx* c, int32_t n, int32_t first_iter, uint32_t seed,
                   const int* iter_dev, int* launches, float nh_frac = 0.f) {
  if (mbir_sweep_fits(c)) return -1;
  SweepArgs A;
  A.mag = c->nh ? c->d_mag : nullptr;
  A.nh_rank = c->d_vsc;
  A.nh_frac = nh_frac;
  if (nh_frac > 0.f && (c->tile * c->tile > kNhMax || !c->nh))
    return fail("non-homogeneous passes need option nhicd and a tile of at most 256 columns");
  A.col.aval = c->d_aval;
  A.col.aval16 = c->afp16 ? c->d_aval16 : nullptr;
  A.col.ascale = c->afp16 ? c->d_ascale : nullptr;
  A.col.afirst = c->d_afirst;
  A.col.frame_views = c->d_frame_views;
  A.col.w = c->d_w;
  A.col.e = c->d_e;
  A.xcd_check = c->d_xcd;
  A.frame_ptr = c->d_frame_ptr;
  A.tile_ptr = c->d_tile_ptr;
  A.tile_cols = c->d_tile_cols;
  A.x = c->d_x;
  A.stats = c->d_stats;
  A.alpha = (float)(c->alpha > 0 ? c->alpha : c->g.alpha);
  A.tile = c->tile;
  A.tiles_x = (c->g.N_x + c->tile - 1) / c->tile;
  A.nvmax = c->nvmax;
  A.iter_dev = iter_dev;
  const size_t xfl = (size_t)(c->tile + 2) * (c->tile + 2) * c->g.N_z;
  if (xfl * sizeof(float) > kSweepXMax)
    return fail("tile too large for N_z: the tile slab must fit 48 KiB of LDS");
  const size_t xbytes =
      sizeof(float) * sweep_lds_words(xfl, c->nvmax, c->G.W, c->tile * c->tile);
  // static LDS of k_sweep (t+-1 buffers, row sums, per-voxel arrays) < 8 KiB
  if (xbytes + 8 * 1024 > c->lds_max)
    return fail("tile slab and A windows exceed the LDS of one workgroup");
  const int nf = c->g.N_time;
  const bool timed = iter_dev == nullptr;
  // register-light prefetch for frames of at most 64 views and 256 A values
  const bool small = c->nvmax <= kWave && c->nvmax * c->G.W <= 4 * kWave;
  *launches = 0;
  // the equit's (colour, frame parity) sets in launch order: colour-major,
  // or with pair_parities (placement 0) (colour k, even frames) together
  // with (colour k + 1, odd frames) -- no voxel of one set neighbours a
  // voxel of the other (other frames, other columns) and their frames own
  // disjoint views, so the two run in one launch
  struct Set {
    int colour, fpar;
  };
  std::vector<std::pair<Set, Set>> plan;
  const bool pair = c->pair && c->placement == 0;
  for (int k = 0; k < 4; ++k) {
    if (pair) {
      plan.push_back({{k, 0}, {(k + 1) % 4, 1}});
    } else {
      plan.push_back({{k, 0}, {-1, 0}});
      plan.push_back({{k, 1}, {-1, 0}});
    }
  }
  auto set_ntl = [&](const Set& q) {
    return q.colour < 0 ? 0 : c->launch_off[q.colour + 1] - c->launch_off[q.colour];
  };
  auto set_frames = [&](const Set& q) { return (nf - q.fpar + 1) / 2; };
  for (int it = 0; it < n; ++it) {
    A.iter = (timed ? first_iter : 0) + it;
    for (const auto& pr : plan) {
      Set a = pr.first, b = pr.second;
      bool ha = set_ntl(a) > 0 && set_frames(a) > 0, hb = set_ntl(b) > 0 && set_frames(b) > 0;
      if (!ha && !hb) continue;
      if (!ha) {
        a = b;
        hb = false;
      }
      {
        const int colour = a.colour, fpar = a.fpar, ntl = set_ntl(a), frames = set_frames(a);
        A.launch_tiles = c->d_launch_tiles + c->launch_off[colour];
        A.n_launch_tiles = ntl;
        A.fparity = fpar;
        A.seed = seed * 2654435761u + (unsigned)(it * 8 + colour * 2 + fpar) * 40503u;
        unsigned nwg = (unsigned)(ntl * frames);
        A.place_r = 0;
        A.nfl = frames;
        A.unit_slots = 0;
        if (c->placement == 1) {
          // frames bound to XCDs: r parts per frame, lcm(frames, 8) units
          int ga = frames, gb = kXcds;
          while (gb) {
            const int t = ga % gb;
            ga = gb;
            gb = t;
          }
          A.place_r = kXcds / ga;
          A.unit_slots = (ntl + A.place_r - 1) / A.place_r;
          nwg = (unsigned)(frames * A.place_r * A.unit_slots);
        }
        A.nwg1 = 0x7fffffff;
        A.launch_tiles2 = nullptr;
        A.n_launch_tiles2 = 1;
        A.fparity2 = 0;
        A.seed2 = 0;
        if (hb) {
          A.nwg1 = (int)nwg;
          A.launch_tiles2 = c->d_launch_tiles + c->launch_off[b.colour];
          A.n_launch_tiles2 = set_ntl(b);
          A.fparity2 = b.fpar;
          A.seed2 = seed * 2654435761u + (unsigned)(it * 8 + b.colour * 2 + b.fpar) * 40503u;
          nwg += (unsigned)(set_ntl(b) * set_frames(b));
        }
        if (timed && ev_begin(c, MBIR_K_SWEEP)) return -1;
        const dim3 grid(nwg), block(kSweepThreads);
        if (small && c->afp16) {
          MBIR_DISPATCH_W(c->G.W, hipLaunchKernelGGL((k_sweep<kW, true, true>), grid, block,
                                                     xbytes, c->stream, c->SG, A));
        } else if (small) {
          MBIR_DISPATCH_W(c->G.W, hipLaunchKernelGGL((k_sweep<kW, true, false>), grid, block,
                                                     xbytes, c->stream, c->SG, A));
        } else if (c->afp16) {
          MBIR_DISPATCH_W(c->G.W, hipLaunchKernelGGL((k_sweep<kW, false, true>), grid, block,
                                                     xbytes, c->stream, c->SG, A));
        } else {
          MBIR_DISPATCH_W(c->G.W, hipLaunchKernelGGL((k_sweep<kW, false, false>), grid, block,
                                                     xbytes, c->stream, c->SG, A));
        }
        if (timed) {
          if (ev_end(c, MBIR_K_SWEEP)) return -1;
        } else {
          HIP_OK(hipGetLastError());
        }
        ++*launches;
      }
    }
  }
  return 0;
}

}  // namespace

int mbir_hip_icd_equits(mbir_hip_ctx* c, int32_t n, int32_t first_iter, uint32_t seed,
                        double* stats) {
  if (!c) return fail("null context");
  if (!c->a_built) return fail("A not built");
  if (build_tiles(c)) return -1;
  HIP_OK(hipMemsetAsync(c->d_stats, 0, 4 * sizeof(double), c->stream));
  int launches = 0;
  if (sweep_launches(c, n, first_iter, seed, nullptr, &launches)) return -1;
  if (stats) {
    HIP_OK(hipMemcpyAsync(stats, c->d_stats, 4 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

namespace {

// Windows of the rectangular-patch offset constraint: triangular weights
// over blocks of 2 N_r / round(sqrt N_r) channels by 2 N_t / round(sqrt N_t)
// rows, half-overlapping, wrapped around both axes, a wrapped sample
// overwriting (XT_OffsetError.c:46-96).  Returns the number of windows.
int offset_windows(int N_r, int N_t, std::vector<double>* win) {
  const int rs = 2 * N_r / (int)(std::sqrt((double)N_r) + 0.5);
  const int ts = 2 * N_t / (int)(std::sqrt((double)N_t) + 0.5);
  if (rs / 2 < 1 || ts / 2 < 1) return 0;
  int m = 0;
  for (int i = 0; i <= N_r - rs / 2; i += rs / 2)
    for (int j = 0; j <= N_t - ts / 2; j += ts / 2) {
      if (win) {
        win->resize((size_t)(m + 1) * N_r * N_t, 0.0);
        double* wm = win->data() + (size_t)m * N_r * N_t;
        for (int k = i; k < i + rs; ++k)
          for (int l = j; l < j + ts; ++l)
            wm[(size_t)(k % N_r) * N_t + (l % N_t)] =
                ((k - i) < rs / 2 ? (k - i + 1) : rs - (k - i)) *
                (double)((l - j) < ts / 2 ? (l - j + 1) : ts - (l - j));
      }
      ++m;
    }
  return m;
}

// a := a^-1 (n x n, Gauss-Jordan with partial pivoting); false if singular
bool invert_in_place(std::vector<double>& a, int n) {
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (a[(size_t)piv * n + c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double p = a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] /= p;
      inv[(size_t)c * n + k] /= p;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  a.swap(inv);
  return true;
}

}  // namespace

int mbir_hip_offset_update(mbir_hip_ctx* c, int32_t mode) {
  if (!c) return fail("null context");
  if (mode < 1 || mode > 3) return fail("offset mode must be 1, 2 or 3");
  const mbir_hip_geom& g = c->g;
  const int nrt = g.N_r * g.N_t;
  if (!c->d_d) {
    if (dev_alloc(&c->d_d, (size_t)nrt)) return -1;
    HIP_OK(hipMemsetAsync(c->d_d, 0, (size_t)nrt * sizeof(float), c->stream));
  }
  if (!c->d_rt && dev_alloc(&c->d_rt, (size_t)3 * nrt)) return -1;
  hipLaunchKernelGGL(k_offset_sums, dim3((nrt + 255) / 256), dim3(256), 0, c->stream, g.N_p, nrt,
                     g.huber_T, g.huber_delta, c->d_e, c->d_w, c->d_d, mode == 3 ? 1 : 0,
                     c->d_rt, c->d_rt + nrt);
  HIP_OK(hipGetLastError());
  std::vector<double> sums((size_t)2 * nrt), step(nrt);
  HIP_OK(hipMemcpyAsync(sums.data(), c->d_rt, sums.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  const double* num = sums.data();
  const double* lam = sums.data() + nrt;
  std::vector<float>& d = c->offset;
  if (mode == 1) {  // update_d_offset_unconstrained, XT_ICD_update.c:866-911
    for (int i = 0; i < nrt; ++i) step[i] = num[i] / lam[i];
  } else if (mode == 2) {  // zero-mean offset, XT_ICD_update.c:813-863
    double ns = 0.0, ds = 0.0;
    for (int i = 0; i < nrt; ++i) {
      ns += d[i] + num[i] / lam[i];
      ds += 1.0 / lam[i];
    }
    const double gamma = ns / ds;
    for (int i = 0; i < nrt; ++i) step[i] = (num[i] - gamma) / lam[i];
  } else {  // windowed zero-sum constraints, XT_ICD_update.c:749-810
    // The windows overlap only their neighbours, so every sum runs over the
    // windows covering each entry (cached: they depend on the detector
    // only), in the same order as the full sums over all entries and
    // windows, whose other terms are zero.
    if (c->off_cover_ptr.empty()) {
      std::vector<double> A;
      c->off_windows = offset_windows(g.N_r, g.N_t, &A);
      c->off_cover_ptr.assign((size_t)nrt + 1, 0);
      for (int k = 0; k < nrt; ++k) {
        for (int m = 0; m < c->off_windows; ++m) {
          const double a = A[(size_t)m * nrt + k];
          if (a != 0.0) {
            c->off_cover_win.push_back(m);
            c->off_cover_val.push_back(a);
          }
        }
        c->off_cover_ptr[k + 1] = (int)c->off_cover_win.size();
      }
    }
    const int M = c->off_windows;
    if (M == 0) return fail("detector too small for the offset windows");
    const std::vector<int>& cp = c->off_cover_ptr;
    const std::vector<int>& cw = c->off_cover_win;
    const std::vector<double>& cv = c->off_cover_val;
    std::vector<double> D((size_t)M * M, 0.0), v(M, 0.0), cm(M);
    for (int k = 0; k < nrt; ++k)
      for (int p = cp[k]; p < cp[k + 1]; ++p) {
        v[cw[p]] += cv[p] * num[k] / lam[k];
        for (int q = cp[k]; q < cp[k + 1]; ++q)
          D[(size_t)cw[p] * M + cw[q]] += cv[p] * cv[q] / lam[k];
      }
    if (!invert_in_place(D, M)) return fail("offset constraint matrix is singular");
    for (int i = 0; i < M; ++i) {
      double s = 0.0;
      for (int j = 0; j < M; ++j) s += D[(size_t)i * M + j] * v[j];
      cm[i] = s;
    }
    for (int k = 0; k < nrt; ++k) {
      double s = 0.0;
      for (int p = cp[k]; p < cp[k + 1]; ++p) s += cv[p] * cm[cw[p]];
      step[k] = (num[k] - s) / lam[k] - d[k];  // new offset minus old
    }
  }
  for (int i = 0; i < nrt; ++i) d[i] = (float)((double)d[i] + step[i]);
  HIP_OK(hipMemcpyAsync(c->d_rt + 2 * nrt, step.data(), (size_t)nrt * sizeof(double),
                        hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->d_d, d.data(), (size_t)nrt * sizeof(float), hipMemcpyHostToDevice,
                        c->stream));
  const long long n = (long long)c->nsino;
  hipLaunchKernelGGL(k_offset_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     n, nrt, c->d_rt + 2 * nrt, c->d_e);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_get_offset(mbir_hip_ctx* c, float* d) {
  if (!c || !d) return fail("null argument");
  memcpy(d, c->offset.data(), c->offset.size() * sizeof(float));
  return 0;
}

int mbir_hip_variance_update(mbir_hip_ctx* c, double* var_est) {
  if (!c) return fail("null context");
  const long long n = (long long)c->nsino;
  const unsigned nb = 1024;
  if (ensure_partial(c, nb)) return -1;
  const double var = c->g.var_est;
  hipLaunchKernelGGL(k_var_sums, dim3(nb), dim3(256), 0, c->stream, n, c->g.huber_T,
                     c->g.huber_delta, var, c->d_e, c->d_w, c->d_partial);
  double s = 0.0;
  if (sum_partials(c, nb, &s)) return -1;
  const double vnew = s / (double)n;
  if (!(vnew > 0.0)) return fail("variance estimate is not positive");
  hipLaunchKernelGGL(k_var_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                     var, vnew, c->d_w);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  c->g.var_est = vnew;
  if (var_est) *var_est = vnew;
  return 0;
}

int mbir_hip_get_weight(mbir_hip_ctx* c, float* w) {
  if (!c || !w) return fail("null argument");
  HIP_OK(hipMemcpy(w, c->d_w, c->nsino * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int mbir_hip_nhicd(mbir_hip_ctx* c, int32_t passes, double fraction, int32_t iter,
                   uint32_t seed, double* stats) {
  if (!c) return fail("null context");
  if (!c->a_built) return fail("A not built");
  if (!c->nh) return fail("option nhicd is off: no magnitude map");
  if (passes < 0 || !(fraction > 0.0 && fraction <= 1.0))
    return fail("passes must be >= 0 and fraction in (0, 1]");
  if (build_tiles(c)) return -1;
  HIP_OK(hipMemsetAsync(c->d_stats, 0, 4 * sizeof(double), c->stream));
  const long long nm = (long long)c->ncol * c->g.N_time;
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_vsc, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, c->stream,
                       c->g.N_time, c->g.N_y, c->g.N_x, c->d_mag, c->d_vsc);
    int launches = 0;
    if (sweep_launches(c, 1, iter, seed + 7919u * (unsigned)(p + 1), nullptr, &launches,
                       (float)fraction))
      return -1;
  }
  HIP_OK(hipGetLastError());
  if (stats) {
    HIP_OK(hipMemcpyAsync(stats, c->d_stats, 4 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int mbir_hip_get_magnitudes(mbir_hip_ctx* c, int32_t filtered, float* out) {
  if (!c || !out) return fail("null argument");
  if (!c->nh) return fail("option nhicd is off: no magnitude map");
  const long long nm = (long long)c->ncol * c->g.N_time;
  if (filtered)
    hipLaunchKernelGGL(k_vsc, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, c->stream,
                       c->g.N_time, c->g.N_y, c->g.N_x, c->d_mag, c->d_vsc);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, filtered ? c->d_vsc : c->d_mag, nm * sizeof(float),
                        hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_icd_graph(mbir_hip_ctx* c, int32_t n, int32_t first_iter, uint32_t seed) {
  if (!c) return fail("null context");
  if (!c->a_built) return fail("A not built");
  if (n < 1) return fail("n must be positive");
  if (build_tiles(c)) return -1;
  if (!c->d_iter && dev_alloc(&c->d_iter, 1)) return -1;
  const long long key[7] = {n, (long long)seed, c->tile, c->afp16 ? 1 : 0,
                            (long long)((c->alpha > 0 ? c->alpha : c->g.alpha) * 1e9),
                            (c->nh ? 1 : 0) | (c->placement << 1) | (c->d_xcd ? 4 : 0) |
                                (c->pair ? 8 : 0),
                            (long long)(uintptr_t)c->d_xcd};
  if (!c->gexec || memcmp(key, c->gkey, sizeof(key)) != 0) {
    if (c->gexec) HIP_OK(hipGraphExecDestroy(c->gexec));
    if (c->graph) HIP_OK(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, c->stream, c->d_stats, 4);
    int launches = 0;
    const int rc = sweep_launches(c, n, 0, seed, c->d_iter, &launches);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (rc || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? -1 : fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    }
    c->graph = g;
    HIP_OK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    memcpy(c->gkey, key, sizeof(key));
    c->glaunches = launches;
  }
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->d_iter, first_iter);
  if (ev_begin(c, MBIR_K_SWEEP)) return -1;
  HIP_OK(hipGraphLaunch(c->gexec, c->stream));
  if (ev_end(c, MBIR_K_SWEEP)) return -1;
  c->launches[MBIR_K_SWEEP] += c->glaunches - 1;  // the graph's launches, not one
  return 0;
}

int mbir_hip_cost(mbir_hip_ctx* c, int32_t fresh, double* out) {
  if (!c || !out) return fail("null argument");
  const mbir_hip_geom& g = c->g;
  double data = 0.0, prior = 0.0;
  if (fresh < 0 || fresh > 2) return fail("fresh must be 0, 1 or 2");
  if (fresh == 1) {
    if (run_fwdproj(c, 1, c->d_x, nullptr)) return -1;
    const size_t nch = (g.N_t + fwd_tchunk(g, 8) - 1) / fwd_tchunk(g, 8);
    if (sum_partials(c, (size_t)g.N_p * nch, &data)) return -1;
  } else {
    const unsigned blocks = 1024;
    if (ensure_partial(c, blocks)) return -1;
    if (ev_begin(c, MBIR_K_COST)) return -1;
    hipLaunchKernelGGL(k_data_cost, dim3(blocks), dim3(256), 0, c->stream, c->G, c->nsino,
                       c->d_e, c->d_w, c->d_partial);
    if (ev_end(c, MBIR_K_COST)) return -1;
    if (sum_partials(c, blocks, &data)) return -1;
  }
  data /= 2.0;
  data += (double)g.N_p * (double)g.N_r * (double)g.N_t * log(g.var_est) / 2;
  {
    if ((long long)g.N_y * g.N_x * g.N_z * g.N_time >= 0x7fffffffLL)
      return fail("too many voxels for the prior cost");
    const unsigned blocks = 1024;
    if (ensure_partial(c, blocks)) return -1;
    if (ev_begin(c, MBIR_K_COST)) return -1;
    if (fresh == 2)
      hipLaunchKernelGGL(k_prior_cost<true>, dim3(blocks), dim3(256), 0, c->stream, c->G,
                         c->d_x, c->d_partial);
    else
      hipLaunchKernelGGL(k_prior_cost<false>, dim3(blocks), dim3(256), 0, c->stream, c->G,
                         c->d_x, c->d_partial);
    if (ev_end(c, MBIR_K_COST)) return -1;
    if (sum_partials(c, blocks, &prior)) return -1;
  }
  out[0] = data + prior;
  out[1] = data;
  out[2] = prior;
  return 0;
}

int mbir_hip_substitution_batch(const mbir_hip_geom* g, int32_t n, const float* V,
                                const float* theta1, const float* theta2, const float* nb,
                                const uint8_t* nbf, const float* tn, const uint8_t* tf,
                                float* out) {
  if (!g || n < 0 || (n > 0 && (!V || !theta1 || !theta2 || !nb || !nbf || !tn || !tf || !out)))
    return fail("bad argument");
  if (n == 0) return 0;
  if (g->q != 2.0) return fail("q must be 2 (the qGGMRF of the sweep kernels)");
  const SwGeom G = make_sw_geom(*g, 0);
  float *dV, *d1, *d2, *dnb, *dtn, *dout;
  unsigned char *dnbf, *dtf;
  if (dev_alloc(&dV, n) || dev_alloc(&d1, n) || dev_alloc(&d2, n) ||
      dev_alloc(&dnb, (size_t)n * 27) || dev_alloc(&dtn, (size_t)n * 2) || dev_alloc(&dout, n) ||
      dev_alloc(&dnbf, (size_t)n * 27) || dev_alloc(&dtf, (size_t)n * 2))
    return -1;
  HIP_OK(hipMemcpy(dV, V, n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d1, theta1, n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d2, theta2, n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dnb, nb, (size_t)n * 27 * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dtn, tn, (size_t)n * 2 * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dnbf, nbf, (size_t)n * 27, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dtf, tf, (size_t)n * 2, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_subst, dim3((unsigned)((n + 3) / 4)), dim3(kWave), 0, 0, G, (float)g->alpha, n, dV, d1, d2,
                     dnb, dnbf, dtn, dtf, dout);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
  void* bufs[] = {dV, d1, d2, dnb, dtn, dout, dnbf, dtf};
  for (void* b : bufs) HIP_OK(hipFree(b));
  return 0;
}

}  // extern "C"

// Device scratch of mbir_hip_substitution_batch_noalloc: the staging copies
// of its inputs and output, in one array per GPU that the library holds
// from module load until the process ends (no allocation per call).
constexpr size_t kNoAllocBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (hipGetSymbolAddress, once per
// device with that device current); host code uses only this address.
int scratch_address(void** out) {
  int dev = -1;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("hipGetSymbolAddress returned a null address for scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the staged arrays of n voxels in the scratch array
// (V, theta1, theta2, nb, tn, out as floats, then nbf, tf as bytes); returns
// the total.
size_t noalloc_layout(size_t n, size_t off[8]) {
  const size_t sz[8] = {4 * n, 4 * n, 4 * n, 4 * 27 * n, 4 * 2 * n, 4 * n, 27 * n, 2 * n};
  size_t at = 0;
  for (int k = 0; k < 8; ++k) {
    off[k] = at;
    at += align256(sz[k]);
  }
  return at;
}

}  // namespace

extern "C" {

int mbir_hip_substitution_batch_noalloc(const mbir_hip_geom* g, int32_t n, const float* V,
                                        const float* theta1, const float* theta2,
                                        const float* nb, const uint8_t* nbf, const float* tn,
                                        const uint8_t* tf, float* out) {
  if (!g || n < 0 || (n > 0 && (!V || !theta1 || !theta2 || !nb || !nbf || !tn || !tf || !out)))
    return fail("bad argument");
  if (n > MBIR_SUBST_NOALLOC_MAX)
    return fail("n exceeds MBIR_SUBST_NOALLOC_MAX (" + std::to_string(MBIR_SUBST_NOALLOC_MAX) +
                " voxels) of the no-allocation substitution");
  if (n == 0) return 0;
  if (g->q != 2.0) return fail("q must be 2 (the qGGMRF of the sweep kernels)");
  size_t off[8];
  if (noalloc_layout((size_t)n, off) > kNoAllocBytes)
    return fail("staging layout exceeds the scratch array");
  unsigned char* base = nullptr;
  if (scratch_address((void**)&base)) return -1;
  float* dV = (float*)(base + off[0]);
  float* d1 = (float*)(base + off[1]);
  float* d2 = (float*)(base + off[2]);
  float* dnb = (float*)(base + off[3]);
  float* dtn = (float*)(base + off[4]);
  float* dout = (float*)(base + off[5]);
  unsigned char* dnbf = base + off[6];
  unsigned char* dtf = base + off[7];
  const SwGeom G = make_sw_geom(*g, 0);
  HIP_OK(hipMemcpy(dV, V, (size_t)n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d1, theta1, (size_t)n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d2, theta2, (size_t)n * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dnb, nb, (size_t)n * 27 * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dtn, tn, (size_t)n * 2 * 4, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dnbf, nbf, (size_t)n * 27, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dtf, tf, (size_t)n * 2, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_subst, dim3((unsigned)((n + 3) / 4)), dim3(kWave), 0, 0, G,
                     (float)g->alpha, n, dV, d1, d2, dnb, dnbf, dtn, dtf, dout);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int mbir_hip_voxel_thetas(mbir_hip_ctx* c, int32_t n, const int32_t* tzyx, float* theta1,
                          float* theta2) {
  if (!c || n < 0 || (n > 0 && (!tzyx || !theta1 || !theta2))) return fail("bad argument");
  if (!c->a_built) return fail("A not built");
  if (mbir_sweep_fits(c)) return -1;
  if (n == 0) return 0;
  for (int i = 0; i < n; ++i) {
    const int32_t* q = tzyx + 4 * i;
    if (q[0] < 0 || q[0] >= c->g.N_time || q[1] < 0 || q[1] >= c->g.N_z || q[2] < 0 ||
        q[2] >= c->g.N_y || q[3] < 0 || q[3] >= c->g.N_x)
      return fail("voxel index out of range");
  }
  int* dq;
  float *d1, *d2;
  if (dev_alloc(&dq, (size_t)n * 4) || dev_alloc(&d1, n) || dev_alloc(&d2, n)) return -1;
  HIP_OK(hipMemcpyAsync(dq, tzyx, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
  ColumnArgs A;
  A.aval = c->d_aval;  // the FP32 copy (FP16-rounded values when afp16 is on)
  A.aval16 = nullptr;
  A.ascale = nullptr;
  A.afirst = c->d_afirst;
  A.frame_views = c->d_frame_views;
  A.w = c->d_w;
  A.e = c->d_e;
  const size_t shm = sizeof(float) * sweep_lds_words(0, c->nvmax, c->G.W, 0);
  MBIR_DISPATCH_W(c->G.W, hipLaunchKernelGGL(k_thetas<kW>, dim3(n), dim3(kWave), shm, c->stream,
                                             c->SG, A, c->d_frame_ptr, dq, c->nvmax, d1, d2));
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(theta1, d1, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(theta2, d2, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipFree(dq));
  HIP_OK(hipFree(d1));
  HIP_OK(hipFree(d2));
  return 0;
}

int mbir_hip_pc_enable(mbir_hip_ctx* c, const mbir_hip_pc_prior* pr) {
  if (!c || !pr) return fail("null argument");
  if (pr->q != 2.0) return fail("q must be 2 (the qGGMRF of the tomographic kernels)");
  if (pr->mu <= 0 || pr->wavenumber <= 0) return fail("mu and the wavenumber must be positive");
  if (c->g.N_z > kNZPc) return fail("the tomographic sweep takes at most 256 voxels per column");
  if (!c->d_x2 && (dev_alloc(&c->d_x2, c->nvox) || dev_alloc(&c->d_e2, c->nsino))) return -1;
  HIP_OK(hipMemset(c->d_x2, 0, c->nvox * sizeof(float)));
  HIP_OK(hipMemset(c->d_e2, 0, c->nsino * sizeof(float)));
  PcTab& P = c->ptab;
  PcDbl& D = c->pdbl;
  memset(&P, 0, sizeof(P));
  memset(&D, 0, sizeof(D));
  for (int i = 0; i < 4; ++i) P.T[i] = (float)(D.T[i] = pr->T[i]);
  P.qmp = (float)(pr->q - pr->p);
  D.p = pr->p;
  D.q = pr->q;
  for (int o = 0; o < 2; ++o) {
    const double cs[2] = {pr->c_s[o], pr->c_t[o]}, sq[2] = {pr->sig_s_q[o], pr->sig_t_q[o]};
    const double sqp[2] = {pr->sig_s_qp[o], pr->sig_t_qp[o]};
    for (int kind = 0; kind < 2; ++kind) {
      P.c[kind][o] = (float)cs[kind];
      P.inv_sq[kind][o] = (float)(1.0 / sq[kind]);
      P.inv_sqp[kind][o] = (float)(1.0 / sqp[kind]);
      P.b0[kind][o] = (float)(pr->q / (sq[kind] * cs[kind]));
      D.c[kind][o] = cs[kind];
      D.sq[kind][o] = sq[kind];
      D.sqp[kind][o] = sqp[kind];
    }
  }
  for (int k = 0; k < 27; ++k) P.filt[k] = (float)(D.filt[k] = pr->filt[k]);
  P.filt[27] = (float)(D.filt[27] = pr->filt_t);
  P.mu = (float)(D.mu = pr->mu);
  P.kw = (float)pr->wavenumber;
  P.alpha = (float)pr->alpha;
  c->pc = true;
  return 0;
}

int mbir_hip_pc_set_objects(mbir_hip_ctx* c, const float* delta, const float* beta) {
  if (!c || !delta || !beta) return fail("null argument");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  if (mbir_hip_set_object(c, delta)) return -1;
  HIP_OK(hipMemcpyAsync(c->d_stage, beta, c->nvox * sizeof(float), hipMemcpyHostToDevice,
                        c->stream));
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  hipLaunchKernelGGL(k_to_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_stage, c->d_x2);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_pc_get_objects(mbir_hip_ctx* c, float* delta, float* beta) {
  if (!c || !delta || !beta) return fail("null argument");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  if (mbir_hip_get_object(c, delta)) return -1;
  const unsigned blocks = (unsigned)((c->nvox + 255) / 256);
  hipLaunchKernelGGL(k_from_colmajor, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_z, c->ncol, c->d_x2, c->d_stage);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(beta, c->d_stage, c->nvox * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_pc_init_error(mbir_hip_ctx* c, const float* target_delta, const float* target_beta) {
  if (!c || !target_delta || !target_beta) return fail("null argument");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  if (!c->a_built) return fail("A not built");
  const double k = (double)c->ptab.kw;
  const float* tg[2] = {target_delta, target_beta};
  const float* xs[2] = {c->d_x, c->d_x2};
  float* es[2] = {c->d_e, c->d_e2};
  for (int o = 0; o < 2; ++o) {
    HIP_OK(hipMemcpyAsync(c->d_stage2, tg[o], c->nsino * sizeof(float), hipMemcpyHostToDevice,
                          c->stream));
    if (run_fwdproj(c, 0, xs[o], es[o], c->d_stage2, k)) return -1;
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_pc_get_error(mbir_hip_ctx* c, float* e_delta, float* e_beta) {
  if (!c || !e_delta || !e_beta) return fail("null argument");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  HIP_OK(hipMemcpyAsync(e_delta, c->d_e, c->nsino * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipMemcpyAsync(e_beta, c->d_e2, c->nsino * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_pc_equits(mbir_hip_ctx* c, int32_t n, int32_t first_iter, uint32_t seed,
                       double* stats) {
  if (!c) return fail("null context");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  if (!c->a_built) return fail("A not built");
  if (c->afp16) return fail("the tomographic sweep reads the FP32 A");
  if (build_tiles(c)) return -1;
  const size_t nx = (size_t)c->tile * c->tile * c->g.N_z;
  const size_t words = 2 * nx + (size_t)c->nvmax * (2 + c->G.W) + (size_t)c->tile * c->tile;
  if (words * 4 + 4 * 1024 + 8 * kNZPc * 4 > c->lds_max)
    return fail("tile slabs of both objects exceed the LDS of one workgroup");
  HIP_OK(hipMemsetAsync(c->d_stats, 0, 4 * sizeof(double), c->stream));
  // a wavefront per 64 voxels of a column, at most four
  const int threads = std::min(4, (c->g.N_z + kWave - 1) / kWave) * kWave;
  PcArgs A;
  A.aval = c->d_aval;
  A.afirst = c->d_afirst;
  A.frame_views = c->d_frame_views;
  A.frame_ptr = c->d_frame_ptr;
  A.tile_ptr = c->d_tile_ptr;
  A.tile_cols = c->d_tile_cols;
  A.tiles_x = (c->g.N_x + c->tile - 1) / c->tile;
  A.tile = c->tile;
  A.nvmax = c->nvmax;
  A.W = c->G.W;
  A.xd = c->d_x;
  A.xb = c->d_x2;
  A.ed = c->d_e;
  A.eb = c->d_e2;
  A.stats = c->d_stats;
  const int nf = c->g.N_time;
  for (int it = 0; it < n; ++it) {
    A.iter = first_iter + it;
    for (int colour = 0; colour < 4; ++colour) {
      const int ntl = c->launch_off[colour + 1] - c->launch_off[colour];
      if (ntl == 0) continue;
      A.launch_tiles = c->d_launch_tiles + c->launch_off[colour];
      A.n_launch_tiles = ntl;
      for (int fpar = 0; fpar < 2; ++fpar) {
        const int frames = (nf - fpar + 1) / 2;
        if (frames <= 0) continue;
        A.fparity = fpar;
        A.seed = seed * 2654435761u + (unsigned)(it * 8 + colour * 2 + fpar) * 40503u;
        if (ev_begin(c, MBIR_K_SWEEP)) return -1;
        hipLaunchKernelGGL(k_pc_sweep, dim3((unsigned)(ntl * frames)), dim3(threads),
                           words * sizeof(float), c->stream, c->SG, c->ptab, A);
        if (ev_end(c, MBIR_K_SWEEP)) return -1;
      }
    }
  }
  if (stats) {
    HIP_OK(hipMemcpyAsync(stats, c->d_stats, 3 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_pc_admm_init(mbir_hip_ctx* c, const double* z_delta, const double* z_beta,
                          int32_t from_object) {
  if (!c || !z_delta || !z_beta) return fail("null argument");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  if (!c->a_built) return fail("A not built");
  const size_t n = c->nsino;
  if (!c->d_pz && (dev_alloc(&c->d_pz, 2 * n) || dev_alloc(&c->d_pv, 2 * n) ||
                   dev_alloc(&c->d_pp, 2 * n) || dev_alloc(&c->d_pk, n) ||
                   dev_alloc(&c->d_pdbl, n)))
    return -1;
  const double k = (double)c->ptab.kw;
  const double* zs[2] = {z_delta, z_beta};
  const float* xs[2] = {c->d_x, c->d_x2};
  float* es[2] = {c->d_e, c->d_e2};
  const unsigned blocks = 1024;
  HIP_OK(hipMemsetAsync(c->d_stage2, 0, n * sizeof(float), c->stream));
  for (int o = 0; o < 2; ++o) {
    if (run_fwdproj(c, 0, xs[o], c->d_pk, c->d_stage2, k)) return -1;  // -k A x
    HIP_OK(hipMemcpyAsync(c->d_pdbl, zs[o], n * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    hipLaunchKernelGGL(k_pc_admm_init, dim3(blocks), dim3(256), 0, c->stream, (long long)n,
                       c->d_pdbl, c->d_pk, from_object ? 1 : 0, c->d_pz + o * n,
                       c->d_pv + o * n, c->d_pp + o * n, es[o]);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  c->admm = true;
  return 0;
}

int mbir_hip_pc_admm_target(mbir_hip_ctx* c, double* a_delta, double* a_beta) {
  if (!c || !a_delta || !a_beta) return fail("null argument");
  if (!c->admm) return fail("ADMM state not initialised (mbir_hip_pc_admm_init)");
  const size_t n = c->nsino;
  double* as[2] = {a_delta, a_beta};
  const float* es[2] = {c->d_e, c->d_e2};
  for (int o = 0; o < 2; ++o) {
    hipLaunchKernelGGL(k_pc_admm_target, dim3(1024), dim3(256), 0, c->stream, (long long)n,
                       c->d_pz + o * n, es[o], c->d_pdbl);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(as[o], c->d_pdbl, n * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int mbir_hip_pc_admm_update(mbir_hip_ctx* c, const double* z_delta, const double* z_beta,
                            double* residuals) {
  if (!c || !z_delta || !z_beta) return fail("null argument");
  if (!c->admm) return fail("ADMM state not initialised (mbir_hip_pc_admm_init)");
  const size_t n = c->nsino;
  const unsigned blocks = 1024;
  if (ensure_partial(c, 3 * blocks)) return -1;
  const double* zs[2] = {z_delta, z_beta};
  float* es[2] = {c->d_e, c->d_e2};
  double res[2][2];
  for (int o = 0; o < 2; ++o) {
    HIP_OK(hipMemcpyAsync(c->d_pdbl, zs[o], n * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    hipLaunchKernelGGL(k_pc_admm_update, dim3(blocks), dim3(256), 0, c->stream, (long long)n,
                       c->d_pdbl, c->d_pz + o * n, es[o], c->d_pv + o * n, c->d_pp + o * n,
                       c->d_partial);
    HIP_OK(hipGetLastError());
    std::vector<double> h(3 * blocks);
    HIP_OK(hipMemcpyAsync(h.data(), c->d_partial, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    double pr = 0.0, du = 0.0, vs = 0.0;
    for (unsigned b = 0; b < blocks; ++b) {
      pr += h[3 * b];
      du += h[3 * b + 1];
      vs += h[3 * b + 2];
    }
    res[o][0] = sqrt(pr / (double)n);
    res[o][1] = sqrt(du / vs);
  }
  if (residuals) {  // the reference's order: magnitude (beta) first
    residuals[0] = res[1][0];
    residuals[1] = res[1][1];
    residuals[2] = res[0][0];
    residuals[3] = res[0][1];
  }
  return 0;
}

int mbir_hip_pc_admm_state(mbir_hip_ctx* c, int32_t which, float* delta, float* beta) {
  if (!c || !delta || !beta) return fail("null argument");
  if (!c->admm) return fail("ADMM state not initialised (mbir_hip_pc_admm_init)");
  if (which < 0 || which > 2) return fail("which must be 0 (Z), 1 (V) or 2 (P)");
  const float* src = which == 0 ? c->d_pz : which == 1 ? c->d_pv : c->d_pp;
  const size_t n = c->nsino;
  HIP_OK(hipMemcpyAsync(delta, src, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(beta, src + n, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_pc_cost(mbir_hip_ctx* c, double* out) {
  if (!c || !out) return fail("null argument");
  if (!c->pc) return fail("tomographic mode not enabled (mbir_hip_pc_enable)");
  const unsigned blocks = 1024;
  double data = 0.0, prior = 0.0;
  if (ensure_partial(c, blocks)) return -1;
  hipLaunchKernelGGL(k_pc_data_cost, dim3(blocks), dim3(256), 0, c->stream,
                     (long long)c->nsino, c->pdbl.mu, c->d_e, c->d_e2, c->d_partial);
  HIP_OK(hipGetLastError());
  if (sum_partials(c, blocks, &data)) return -1;
  hipLaunchKernelGGL(k_pc_prior_cost, dim3(blocks), dim3(256), 0, c->stream, c->g.N_time,
                     c->g.N_y, c->g.N_x, c->g.N_z, c->pdbl, c->d_x, c->d_x2, c->d_partial);
  HIP_OK(hipGetLastError());
  if (sum_partials(c, blocks, &prior)) return -1;
  out[0] = data + prior;
  out[1] = data;
  out[2] = prior;
  return 0;
}

int mbir_hip_kernel_time(mbir_hip_ctx* c, int32_t k, double* ms, int64_t* launches) {
  if (!c || k < 0 || k >= MBIR_K_COUNT) return fail("bad argument");
  if (ev_collect(c)) return -1;
  if (ms) *ms = c->ms[k];
  if (launches) *launches = c->launches[k];
  return 0;
}

int mbir_hip_xcd_check(mbir_hip_ctx* c, int64_t* items, int64_t* off_xcd) {
  if (!c || !items || !off_xcd) return fail("null argument");
  if (!c->d_xcd) return fail("option xcd_check is not set");
  unsigned h[2] = {0, 0};
  HIP_OK(hipMemcpyAsync(h, c->d_xcd, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  *items = h[0];
  *off_xcd = h[1];
  return 0;
}

int mbir_hip_reset_kernel_time(mbir_hip_ctx* c) {
  if (!c) return fail("null context");
  if (ev_collect(c)) return -1;
  for (int k = 0; k < MBIR_K_COUNT; ++k) {
    c->ms[k] = 0;
    c->launches[k] = 0;
  }
  return 0;
}

int mbir_hip_synchronize(mbir_hip_ctx* c) {
  if (!c) return fail("null context");
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int mbir_hip_phantom_project(const float* mag, const float* phase, int32_t n_xy, int32_t n_z,
                             int32_t n_p, const float* angles, int32_t proj_rows,
                             int32_t proj_cols, double vox_wid, double rot_center,
                             double wavenumber, float* proj_real, float* proj_imag) {
  if (!mag || !phase || !angles || !proj_real || !proj_imag) return fail("null argument");
  if (n_xy < 2 || n_z < 1 || n_p < 1 || proj_rows < 1 || proj_cols < 1 || n_xy % proj_cols)
    return fail("phantom sizes must be positive and n_xy a multiple of proj_cols");
  // geometry of initPhantomStructures (PCMBIR_3D/XT_Init.c:611-690): the
  // detector sampled at the phantom's voxel pitch
  const double len_r = vox_wid * proj_cols, len_t = vox_wid * proj_rows;
  const double rc = rot_center * (double)(n_xy / proj_cols);
  mbir_hip_geom g{};
  g.N_r = n_xy;
  g.N_t = n_z;
  g.N_p = n_p;
  g.N_x = g.N_y = n_xy;
  g.N_z = n_z;
  g.N_time = 1;
  g.mult_xy = g.mult_z = 1;
  g.delta_r = len_r / n_xy;
  g.delta_t = len_t / n_z;
  g.R0 = -rc * g.delta_r;
  g.RMax = (n_xy - rc) * g.delta_r;
  g.x0 = g.R0;
  g.y0 = -len_r / 2.0;
  g.delta_xy = g.delta_r;
  g.OffsetR = (g.delta_xy / sqrt(2.0) + g.delta_r / 2.0) / kRespBins;
  const double T0 = -len_t / 2.0, TMax = len_t / 2.0, dz = g.delta_t, dt = g.delta_t;
  const double offT = (dz / 2 + dt / 2) / kRespBins;
  // z-line response (ZLineResponseProfile, PCMBIR_3D/XT_Profile.c:182-227)
  // and the rows each slice feeds (storeVoxelLineResponse, :229-290),
  // inverted to the slices feeding each row
  float ht[kRespBins + 1];
  for (int i = 0; i < kRespBins; ++i) {
    const double c = i * offT;
    double h = (c <= dz / 2 - dt / 2) ? dt : -c + dz / 2 + dt / 2;
    if (h < 0) h = 0;
    ht[i] = (float)(h / dt);
  }
  ht[kRespBins] = 0.f;
  std::vector<int> lz((size_t)n_z * kSimL, -1);
  std::vector<float> lw((size_t)n_z * kSimL, 0.f);
  for (int z = 0; z < n_z; ++z) {
    const double t = ((double)z + 0.5) * dz + T0;
    const double tmin = (t - dz / 2) > T0 ? t - dz / 2 : T0;
    const double tmax = (t + dz / 2) <= TMax ? t + dz / 2 : TMax;
    const int lo = std::max(0, (int)floor((tmin - T0) / dt));
    const int hi = std::min(n_z - 1, (int)floor((tmax - T0) / dt));
    for (int it = lo; it <= hi; ++it) {
      const double d = fabs(((double)it + 0.5) * dt + T0 - t);
      const int k = (int)floor(d / offT);
      if (k >= kRespBins) continue;
      const double v = ((k + 1) * offT - d) / offT * ht[k] + (d - k * offT) / offT * ht[k + 1];
      if (v == 0) continue;
      int q = 0;
      while (q < kSimL && lz[(size_t)it * kSimL + q] >= 0) ++q;
      if (q == kSimL) return fail("more than three slices feed one detector row");
      lz[(size_t)it * kSimL + q] = z;
      lw[(size_t)it * kSimL + q] = (float)v;
    }
  }
  std::vector<float> H((size_t)n_p * kNB);
  std::vector<double> cs(n_p), sn(n_p);
  mbir_detector_response(&g, angles, H.data());
  mbir_sincos(n_p, angles, cs.data(), sn.data());
  DevGeom G{};
  G.N_r = n_xy;
  G.N_x = G.N_y = n_xy;
  G.N_z = n_z;
  G.N_p = n_p;
  G.W = 3;
  G.delta_r = g.delta_r;
  G.R0 = g.R0;
  G.RMax = g.RMax;
  G.x0 = g.x0;
  G.y0 = g.y0;
  G.delta_xy = g.delta_xy;
  G.OffsetR = g.OffsetR;
  const int tchunk = std::max(1, std::min({kSimTc, n_z, (int)(48 * 1024 / (8 * n_xy))}));
  if ((size_t)8 * n_xy * tchunk > 48 * 1024)
    return fail("phantom too wide: one detector row of both projections must fit 48 KiB of LDS");
  const size_t nvox = (size_t)n_z * n_xy * n_xy, nout = (size_t)n_p * n_z * n_xy;
  struct Bufs {
    float *m = nullptr, *ph = nullptr, *h = nullptr, *w = nullptr, *orr = nullptr, *oi = nullptr;
    double *c = nullptr, *s = nullptr;
    int* z = nullptr;
    ~Bufs() {
      for (void* b : {(void*)m, (void*)ph, (void*)h, (void*)w, (void*)orr, (void*)oi, (void*)c,
                      (void*)s, (void*)z})
        if (b) (void)hipFree(b);
    }
  } B;
  if (dev_alloc(&B.m, nvox) || dev_alloc(&B.ph, nvox) || dev_alloc(&B.h, H.size()) ||
      dev_alloc(&B.w, lw.size()) || dev_alloc(&B.z, lz.size()) || dev_alloc(&B.c, cs.size()) ||
      dev_alloc(&B.s, sn.size()) || dev_alloc(&B.orr, nout) || dev_alloc(&B.oi, nout))
    return -1;
  HIP_OK(hipMemcpy(B.m, mag, sizeof(float) * nvox, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.ph, phase, sizeof(float) * nvox, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.h, H.data(), sizeof(float) * H.size(), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.w, lw.data(), sizeof(float) * lw.size(), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.z, lz.data(), sizeof(int) * lz.size(), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.c, cs.data(), sizeof(double) * cs.size(), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.s, sn.data(), sizeof(double) * sn.size(), hipMemcpyHostToDevice));
  const dim3 grid((unsigned)n_p, (unsigned)((n_z + tchunk - 1) / tchunk));
  const size_t lds = sizeof(float) * 2 * (size_t)n_xy * tchunk;
  hipLaunchKernelGGL(k_phantom_project, grid, dim3(256), lds, nullptr, G, B.h, B.c, B.s, B.m,
                     B.ph, B.z, B.w, n_z, tchunk, wavenumber, B.orr, B.oi);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(proj_real, B.orr, sizeof(float) * nout, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(proj_imag, B.oi, sizeof(float) * nout, hipMemcpyDeviceToHost));
  return 0;
}

int mbir_hip_std_downsample(const float* proj, const float* weights_in, int32_t n_proj,
                            int32_t rows, int32_t cols, int32_t ratio, double bh_coef,
                            float* projections, float* weights) {
  if (!proj || !weights_in || !projections || !weights) return fail("null argument");
  if (n_proj < 1 || rows < 1 || cols < 1 || ratio < 1)
    return fail("counts, rows, cols and ratio must be positive");
  const size_t nin = (size_t)n_proj * rows * ratio * cols * ratio;
  const size_t nout = (size_t)n_proj * rows * cols;
  struct Bufs {
    float *p = nullptr, *w = nullptr, *op = nullptr, *ow = nullptr;
    ~Bufs() {
      for (void* b : {(void*)p, (void*)w, (void*)op, (void*)ow})
        if (b) (void)hipFree(b);
    }
  } B;
  if (dev_alloc(&B.p, nin) || dev_alloc(&B.w, nin) || dev_alloc(&B.op, nout) ||
      dev_alloc(&B.ow, nout))
    return -1;
  HIP_OK(hipMemcpy(B.p, proj, sizeof(float) * nin, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.w, weights_in, sizeof(float) * nin, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_std_downsample, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0,
                     nullptr, B.p, B.w, n_proj, rows, cols, ratio, bh_coef, B.op, B.ow);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(projections, B.op, sizeof(float) * nout, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(weights, B.ow, sizeof(float) * nout, hipMemcpyDeviceToHost));
  return 0;
}

int mbir_hip_aps_normalize(const uint16_t* proj, const uint16_t* white, const uint16_t* dark,
                           int32_t n_proj, int32_t n_white, int32_t n_dark, int32_t rows,
                           int32_t cols, int32_t ratio, double bh_coef, float* projections,
                           float* weights) {
  if (!proj || !white || !dark || !projections || !weights) return fail("null argument");
  if (n_proj < 1 || n_white < 1 || n_dark < 1 || rows < 1 || cols < 1 || ratio < 1)
    return fail("counts, rows, cols and ratio must be positive");
  const long long npix = (long long)rows * ratio * cols * ratio;  // detector pixels per frame
  const long long nout = (long long)n_proj * rows * cols;
  struct Bufs {
    uint16_t *p = nullptr, *wh = nullptr, *dk = nullptr;
    float *mw = nullptr, *md = nullptr, *op = nullptr, *ow = nullptr;
    ~Bufs() {
      for (void* b : {(void*)p, (void*)wh, (void*)dk, (void*)mw, (void*)md, (void*)op, (void*)ow})
        if (b) (void)hipFree(b);
    }
  } B;
  if (dev_alloc(&B.p, (size_t)(npix * n_proj)) || dev_alloc(&B.wh, (size_t)(npix * n_white)) ||
      dev_alloc(&B.dk, (size_t)(npix * n_dark)) || dev_alloc(&B.mw, (size_t)npix) ||
      dev_alloc(&B.md, (size_t)npix) || dev_alloc(&B.op, (size_t)nout) ||
      dev_alloc(&B.ow, (size_t)nout))
    return -1;
  HIP_OK(hipMemcpy(B.p, proj, sizeof(uint16_t) * npix * n_proj, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.wh, white, sizeof(uint16_t) * npix * n_white, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.dk, dark, sizeof(uint16_t) * npix * n_dark, hipMemcpyHostToDevice));
  const unsigned bp = (unsigned)((npix + 255) / 256), bo = (unsigned)((nout + 255) / 256);
  hipLaunchKernelGGL(k_aps_mean, dim3(bp), dim3(256), 0, nullptr, B.wh, n_white, npix, B.mw);
  hipLaunchKernelGGL(k_aps_mean, dim3(bp), dim3(256), 0, nullptr, B.dk, n_dark, npix, B.md);
  hipLaunchKernelGGL(k_aps_normalize, dim3(bo), dim3(256), 0, nullptr, B.p, B.mw, B.md, n_proj,
                     rows, cols, ratio, bh_coef, B.op, B.ow);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(projections, B.op, sizeof(float) * nout, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(weights, B.ow, sizeof(float) * nout, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
