// pcmbir_hip.hip -- gfx950 kernels of the PCMBIR phase-retrieval path and the
// C-ABI declared in include/pcmbir_hip.h.
//
// Device layout: complex arrays float2 [n_views][rows][cols], cols fastest.
// Kernels:
//   k_fft_rows_reg   batched FFTs along cols, each sequence in registers
//                    (radix-8 Stockham, LDS between passes, twiddles in LDS)
//   k_fresnel_cols_reg  the column pass likewise, rows/8 threads per column,
//                    up to 64 columns and 512 threads per workgroup (option
//                    fft_col_threads; 8 columns at 512 rows)
//   k_fft_rows       batched in-LDS FFTs along cols; forward = bit-reversed
//                    load + radix-2 DIT, inverse = radix-2 DIF + bit-reversed
//                    store (both unnormalised, FFTW's conventions)
//   k_fresnel_cols   per (view, column tile): FFT along rows, the spectral
//                    factor (Fresnel phase, window, 1/N), inverse FFT along
//                    rows, all in LDS -- one Fresnel transform is three passes
//                    over the data (rows, columns, rows)
//   k_pr_*           elementwise steps and per-view FP64 reductions of the
//                    steepest descent (XT_PhaseRet.c), masked per view
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "pcmbir_hip.h"

namespace {

constexpr int kFftMax = 1024;
constexpr int kThreads = 256;
constexpr int kRowElems = 2048;        // complex values per row-FFT workgroup
constexpr int kColCap = 4096 + 64;     // complex values per column workgroup (padded)

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

#define HIP_OK(call)                                                  \
  do {                                                                \
    hipError_t e_ = (call);                                           \
    if (e_ != hipSuccess)                                             \
      return fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return {a.x * s, a.y * s}; }

__device__ __forceinline__ int brev(int j, int logn) {
  return (int)(__builtin_bitreverse32((unsigned)j) >> (32 - logn));
}

// Radix-2 decimation in time on seq sequences of length n = 2^logn at
// buf[s*stride + i], bit-reversed input, natural-order output;
// tw[m] = exp(-2 pi i m / n), m < n/2.
__device__ void fft_dit(float2* buf, int logn, int seq, int stride, const float2* tw) {
  const int half = 1 << (logn - 1), total = seq * half;
  for (int lh = 0; lh < logn; ++lh) {
    const int h = 1 << lh, tws = logn - 1 - lh;
    __syncthreads();
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int s = t >> (logn - 1), b = t & (half - 1);
      const int k = b & (h - 1);
      const int i0 = s * stride + ((b >> lh) << (lh + 1)) + k, i1 = i0 + h;
      const float2 a = buf[i0], c = cmul(buf[i1], tw[k << tws]);
      buf[i0] = cadd(a, c);
      buf[i1] = csub(a, c);
    }
  }
  __syncthreads();
}

// Radix-2 decimation in frequency with conjugate twiddles (the inverse
// transform, unnormalised): natural-order input, bit-reversed output.
__device__ void fft_dif_inv(float2* buf, int logn, int seq, int stride, const float2* tw) {
  const int half = 1 << (logn - 1), total = seq * half;
  for (int lh = logn - 1; lh >= 0; --lh) {
    const int h = 1 << lh, tws = logn - 1 - lh;
    __syncthreads();
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int s = t >> (logn - 1), b = t & (half - 1);
      const int k = b & (h - 1);
      const int i0 = s * stride + ((b >> lh) << (lh + 1)) + k, i1 = i0 + h;
      const float2 w = tw[k << tws];
      const float2 a = buf[i0], c = buf[i1];
      buf[i0] = cadd(a, c);
      buf[i1] = cmulc(csub(a, c), w);
    }
  }
  __syncthreads();
}

// FFT along cols of nrows contiguous rows of length 2^logn (inv = 0
// forward, 1 inverse), kRowElems values per workgroup.
__global__ __launch_bounds__(kThreads) void k_fft_rows(float2* __restrict__ data, long long nrows,
                                                       int logn, const float2* __restrict__ twg,
                                                       int inv) {
  __shared__ float2 s_buf[kRowElems];
  __shared__ float2 s_tw[kFftMax / 2];
  const int n = 1 << logn;
  const int R = kRowElems >> logn;
  const long long r0 = (long long)blockIdx.x * R;
  const int cnt = (int)((nrows - r0) < R ? (nrows - r0) : R);
  const int tot = cnt << logn;
  float2* src = data + (r0 << logn);
  for (int j = threadIdx.x; j < n / 2; j += blockDim.x) s_tw[j] = twg[j];
  if (!inv) {
    for (int j = threadIdx.x; j < tot; j += blockDim.x)
      s_buf[((j >> logn) << logn) + brev(j & (n - 1), logn)] = src[j];
    fft_dit(s_buf, logn, cnt, n, s_tw);
    for (int j = threadIdx.x; j < tot; j += blockDim.x) src[j] = s_buf[j];
  } else {
    for (int j = threadIdx.x; j < tot; j += blockDim.x) s_buf[j] = src[j];
    fft_dif_inv(s_buf, logn, cnt, n, s_tw);
    for (int j = threadIdx.x; j < tot; j += blockDim.x)
      src[j] = s_buf[((j >> logn) << logn) + brev(j & (n - 1), logn)];
  }
}

// Columns per workgroup of k_fresnel_cols: the largest power of two with
// TC (rows + 1) <= kColCap, at most cols.
__host__ __device__ inline int col_tile(int logr, int logc) {
  int tc = 1;
  while (2 * tc <= (1 << logc) && 2 * tc * ((1 << logr) + 1) <= kColCap) tc <<= 1;
  return tc;
}

// Per (view, tile of TC columns): forward FFT along rows, times fac
// [rows][cols], inverse FFT along rows.  LDS rows are padded by one value
// against bank conflicts.
__global__ __launch_bounds__(kThreads) void k_fresnel_cols(float2* __restrict__ data, int logr,
                                                           int logc,
                                                           const float2* __restrict__ twg,
                                                           const float2* __restrict__ fac) {
  __shared__ float2 s_buf[kColCap];
  __shared__ float2 s_tw[kFftMax / 2];
  const int rows = 1 << logr, cols = 1 << logc;
  const int TC = col_tile(logr, logc), ltc = __builtin_ctz(TC);
  const int stride = rows + 1;
  const int tiles = cols / TC;
  const int view = blockIdx.x / tiles, c0 = (blockIdx.x % tiles) * TC;
  float2* base = data + ((size_t)view << (logr + logc)) + c0;
  for (int j = threadIdx.x; j < rows / 2; j += blockDim.x) s_tw[j] = twg[j];
  const int tot = rows * TC;
  for (int j = threadIdx.x; j < tot; j += blockDim.x) {
    const int r = j >> ltc, c = j & (TC - 1);
    s_buf[c * stride + brev(r, logr)] = base[((size_t)r << logc) + c];
  }
  fft_dit(s_buf, logr, TC, stride, s_tw);
  for (int j = threadIdx.x; j < tot; j += blockDim.x) {
    const int r = j >> ltc, c = j & (TC - 1);
    s_buf[c * stride + r] = cmul(s_buf[c * stride + r], fac[((size_t)r << logc) + c0 + c]);
  }
  fft_dif_inv(s_buf, logr, TC, stride, s_tw);
  for (int j = threadIdx.x; j < tot; j += blockDim.x) {
    const int r = j >> ltc, c = j & (TC - 1);
    base[((size_t)r << logc) + c] = s_buf[c * stride + brev(r, logr)];
  }
}

// ------------------------------------------- register-resident FFTs
// Stockham autosort FFT of length N = 2^logn >= 8 (radix-8 passes, then one
// radix-2 or radix-4 pass for the remaining factor): a sequence is held by
// N/8 threads, thread j holding elements j + (N/8) s, s = 0..7, in
// registers, in natural order on entry and on exit.  A pass of radix R with
// Ns = product of the earlier radices runs 8/R butterflies per thread: the
// virtual thread vt = j + (N/8) q takes elements vt + (N/R) r, multiplies
// them by exp(-+2 pi i (vt mod Ns) r / (Ns R)), transforms them (DFT_R) and
// writes output r at (vt / Ns) Ns R + vt mod Ns + r Ns; between passes the
// sequence goes through LDS (element i at lds[i ls + lo]).  The last pass's
// outputs land where the next read would fetch them, so they stay in
// registers.  inv = 1: exp(+...) (the inverse transform, unnormalised).

// p advanced by a 32-bit byte offset: a wave-uniform base plus a 32-bit
// lane offset, so that an address costs one register, not a 64-bit pair.
template <typename T>
__device__ __forceinline__ T* at_bytes(T* p, unsigned off) {
  using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<C*>(p) + off);
}

// exp(-+2 pi i m / N) from tw[m] = exp(-2 pi i m / N), m < N/2.
template <bool kInv>
__device__ __forceinline__ float2 twid(const float2* __restrict__ tw, int m, int N) {
  const int h = N >> 1;
  float2 w = *at_bytes(tw, (unsigned)(m < h ? m : m - h) * 8u);
  if (m >= h) w = make_float2(-w.x, -w.y);
  if (kInv) w.y = -w.y;
  return w;
}

// times -i (forward) or +i (inverse)
template <bool kInv>
__device__ __forceinline__ float2 rot90(float2 a) {
  return kInv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool kInv>
__device__ __forceinline__ void dft2(float2* a) {
  const float2 t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}

template <bool kInv>
__device__ __forceinline__ void dft4(float2* a) {
  const float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
  const float2 t2 = cadd(a[1], a[3]), t3 = rot90<kInv>(csub(a[1], a[3]));
  a[0] = cadd(t0, t2);
  a[2] = csub(t0, t2);
  a[1] = cadd(t1, t3);
  a[3] = csub(t1, t3);
}

// 8-point DFT: even / odd 4-point DFTs, odd outputs times exp(-+2 pi i k / 8).
template <bool kInv>
__device__ __forceinline__ void dft8(float2* a) {
  float2 e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
  dft4<kInv>(e);
  dft4<kInv>(o);
  const float h = 0.70710678118654752f;  // 1 / sqrt(2)
  o[1] = cmul(o[1], make_float2(h, kInv ? h : -h));
  o[2] = rot90<kInv>(o[2]);
  o[3] = cmul(o[3], make_float2(-h, kInv ? h : -h));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = cadd(e[k], o[k]);
    a[k + 4] = csub(e[k], o[k]);
  }
}

template <int R, bool kInv>
__device__ __forceinline__ void dft_r(float2* a) {
  if constexpr (R == 8) dft8<kInv>(a);
  else if constexpr (R == 4) dft4<kInv>(a);
  else dft2<kInv>(a);
}

template <int R, bool kInv>
__device__ __forceinline__ void stockham_pass(float2 v[8], int j, int logn, int Ns,
                                              const float2* __restrict__ tw, float2* lds, int ls,
                                              int lo, bool last) {
  constexpr int Q = 8 / R;
  const int N = 1 << logn, n8 = N >> 3;
  const int tstep = N / (Ns * R);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int vt = j + n8 * q;
    const int k = vt & (Ns - 1);
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[q + Q * r];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], twid<kInv>(tw, k * tstep * r, N));
    }
    dft_r<R, kInv>(a);
    if (last) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + Q * r] = a[r];
    } else {
      const int base = (vt - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(base + r * Ns) * ls + lo] = a[r];
    }
  }
}

// The whole transform; every thread of the workgroup calls it with the same
// logn (it holds workgroup barriers).
template <bool kInv>
__device__ void fft_reg(float2 v[8], int j, int logn, const float2* __restrict__ tw,
                        float2* lds, int ls, int lo) {
  const int p8 = logn / 3, rem = logn - 3 * p8, n8 = 1 << (logn - 3);
  int Ns = 1;
  for (int p = 0; p < p8; ++p) {
    const bool last = p == p8 - 1 && rem == 0;
    stockham_pass<8, kInv>(v, j, logn, Ns, tw, lds, ls, lo, last);
    Ns <<= 3;
    if (!last) {
      __syncthreads();
#pragma unroll
      for (int s = 0; s < 8; ++s) v[s] = lds[(j + n8 * s) * ls + lo];
      __syncthreads();
    }
  }
  if (rem == 2) stockham_pass<4, kInv>(v, j, logn, Ns, tw, lds, ls, lo, true);
  else if (rem == 1) stockham_pass<2, kInv>(v, j, logn, Ns, tw, lds, ls, lo, true);
}

// Columns per workgroup of k_fresnel_cols_reg (log2): 8 values per thread,
// rows / 8 threads per column, at most max_threads threads (a power of two,
// 64 to 1024), 64 columns and cols, and at most 64 KiB of LDS (the exchange
// buffer of 8 bytes per value and the rows / 2 twiddles).
__host__ __device__ inline int col_tile_reg(int logr, int logc, int max_threads) {
  const auto lds = [logr](int l) { return ((1 << (l + logr)) + (1 << (logr - 1))) * 8; };
  int ltc = 0;
  while (ltc < logc && ltc < 6 && (1 << (ltc + 1 + logr - 3)) <= max_threads &&
         lds(ltc + 1) <= 65536)
    ++ltc;
  return ltc;
}

// Rows >= 8: per (view, tile of TC = 2^ltc columns) the forward FFT along
// rows, times fac [rows][cols], the inverse FFT along rows, each column's
// data in its threads' registers, lanes over the columns of a row (one
// 8 TC-byte segment per row and load); LDS [rows][TC] for the exchanges.
__global__ __launch_bounds__(1024) void k_fresnel_cols_reg(float2* __restrict__ data, int logr,
                                                           int logc, int ltc,
                                                           const float2* __restrict__ twg,
                                                           const float2* __restrict__ fac) {
  extern __shared__ float2 s_ex[];
  const int TC = 1 << ltc, n8 = 1 << (logr - 3);
  const int tiles = 1 << (logc - ltc);
  const int view = blockIdx.x >> (logc - ltc), c0 = (blockIdx.x & (tiles - 1)) << ltc;
  const int col = threadIdx.x & (TC - 1), j = threadIdx.x >> ltc;
  // wave-uniform bases and 32-bit byte offsets (a view is far below 4 GiB)
  float2* vb = data + ((size_t)view << (logr + logc)) + c0;
  const float2* fb = fac + c0;
  const unsigned o0 = ((unsigned)col + ((unsigned)j << logc)) * 8u;
  const unsigned os = ((unsigned)n8 << logc) * 8u;
  float2 v[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = *at_bytes(vb, o0 + s * os);
  // the twiddles in LDS behind the exchange buffer (LDS latency, not an L2
  // round trip, in the passes' dependent twiddle reads)
  float2* s_tw = s_ex + ((size_t)TC << logr);
  for (int i = threadIdx.x; i < (1 << (logr - 1)); i += blockDim.x) s_tw[i] = twg[i];
  __syncthreads();
  fft_reg<false>(v, j, logr, s_tw, s_ex, TC, col);
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = cmul(v[s], *at_bytes(fb, o0 + s * os));
  fft_reg<true>(v, j, logr, s_tw, s_ex, TC, col);
#pragma unroll
  for (int s = 0; s < 8; ++s) *at_bytes(vb, o0 + s * os) = v[s];
}

// Rows of length 2^logn >= 8 (contiguous): 256 threads, 2048 / 2^logn rows
// per workgroup, a row's data in 2^logn / 8 threads' registers.
template <bool kInv>
__global__ __launch_bounds__(256) void k_fft_rows_reg(float2* __restrict__ data, long long nrows,
                                                      int logn,
                                                      const float2* __restrict__ twg) {
  __shared__ float2 s_ex[2048];
  const int n8 = 1 << (logn - 3), S = 256 >> (logn - 3);
  const int seq = threadIdx.x >> (logn - 3), j = threadIdx.x & (n8 - 1);
  const long long row = (long long)blockIdx.x * S + seq;
  const bool live = row < nrows;
  float2* src = data + ((live ? row : 0) << logn) + j;
  __shared__ float2 s_tw[kFftMax / 2];  // the twiddles, read from LDS in the passes
  float2 v[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = live ? src[n8 * s] : make_float2(0.f, 0.f);
  for (int i = threadIdx.x; i < (1 << (logn - 1)); i += 256) s_tw[i] = twg[i];
  __syncthreads();
  fft_reg<kInv>(v, j, logn, s_tw, s_ex, 1, seq << logn);
  if (live) {
#pragma unroll
    for (int s = 0; s < 8; ++s) src[n8 * s] = v[s];
  }
}

// ------------------------------------------------ steepest-descent steps
// All kernels: a grid-stride loop over n_views * npix values, view-masked by
// active[view] (a view that met its stopping test is left untouched).

struct PrArgs {
  const float2 *y, *D, *v;
  const float* lam;
  float2 *w, *om, *g, *b1, *b2;
  const int* active;
  double* scal;  // per view: num, den, alpha, upd
  float nu;
  long long npix, nall;
};

// b = D x
__global__ void k_pr_dmul(PrArgs P, const float2* __restrict__ x, float2* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.nall;
       i += (long long)gridDim.x * blockDim.x) {
    if (!P.active[i / P.npix]) continue;
    b[i] = cmul(P.D[i], x[i]);
  }
}

// b1 = conj(Omega) ((b1 Omega) Lambda)   [mult_OmegaHD then Lambda, then the
// first step of mult_Herm_OmegaHD]; with_y: b2 = conj(Omega) (Lambda y)
__global__ void k_pr_mid(PrArgs P, int with_y) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.nall;
       i += (long long)gridDim.x * blockDim.x) {
    if (!P.active[i / P.npix]) continue;
    const float2 om = P.om[i];
    const float l = P.lam[i];
    const float2 t = cscale(cmul(P.b1[i], om), l);
    P.b1[i] = cmul({om.x, -om.y}, t);
    if (with_y) P.b2[i] = cmul({om.x, -om.y}, cscale(P.y[i], l));
  }
}

// g = b1 conj(D) - b2 conj(D) + nu (w - v)   (compute_gradient, :121-126)
__global__ void k_pr_grad(PrArgs P) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.nall;
       i += (long long)gridDim.x * blockDim.x) {
    if (!P.active[i / P.npix]) continue;
    const float2 d = P.D[i];
    const float2 a = cmulc(P.b1[i], d), b = cmulc(P.b2[i], d);
    const float2 wv = csub(P.w[i], P.v[i]);
    P.g[i] = cadd(csub(a, b), cscale(wv, P.nu));
  }
}

// Block reduction of two FP64 values over the block; result in thread 0.
__device__ void block_sum2(double& a, double& b) {
  __shared__ double ra[kThreads], rb[kThreads];
  __syncthreads();  // a previous call's result has been read by every thread
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      ra[threadIdx.x] += ra[threadIdx.x + s];
      rb[threadIdx.x] += rb[threadIdx.x + s];
    }
    __syncthreads();
  }
  a = ra[0];
  b = rb[0];
}

// The per-view reductions of the step run kPrParts blocks per view, each
// over a contiguous chunk of the view's pixels, into FP64 partial sums
// part[(view kPrParts + block) 2 + {0, 1}]; one thread per view then adds a
// view's partials in block order (deterministic), so a launch has enough
// blocks to fill the chip whatever the number of views.
constexpr int kPrParts = 16;

__device__ __forceinline__ void pr_chunk(long long npix, int b, long long& j0, long long& j1) {
  const long long ch = (npix + kPrParts - 1) / kPrParts;
  j0 = b * ch;
  j1 = j0 + ch < npix ? j0 + ch : npix;
}

// Per view: num = |g|^2, den = Re <g, b1 conj(D)> (compute_stepsize,
// :133-170), partials.  b1 holds H^H(...) of the step's second pass before
// its conj(D) factor.
__global__ __launch_bounds__(kThreads) void k_pr_step(PrArgs P, double* __restrict__ part) {
  const int view = blockIdx.x / kPrParts, blk = blockIdx.x % kPrParts;
  if (!P.active[view]) return;
  const long long o = (long long)view * P.npix;
  long long j0, j1;
  pr_chunk(P.npix, blk, j0, j1);
  double num = 0.0, den = 0.0;
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const float2 g = P.g[o + j];
    const float2 b = cmulc(P.b1[o + j], P.D[o + j]);
    num += (double)g.x * g.x + (double)g.y * g.y;
    den += (double)g.x * b.x + (double)g.y * b.y;
  }
  block_sum2(num, den);
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = num;
    part[2 * (size_t)blockIdx.x + 1] = den;
  }
}

// alpha = num / (den + nu num) per view from the partials.
__global__ void k_pr_step_fin(PrArgs P, const double* __restrict__ part, int n_views) {
  const int view = blockIdx.x * blockDim.x + threadIdx.x;
  if (view >= n_views || !P.active[view]) return;
  double num = 0.0, den = 0.0;
  for (int b = 0; b < kPrParts; ++b) {
    num += part[2 * ((size_t)view * kPrParts + b)];
    den += part[2 * ((size_t)view * kPrParts + b) + 1];
  }
  den += (double)P.nu * num;
  P.scal[4 * view + 0] = num;
  P.scal[4 * view + 1] = den;
  P.scal[4 * view + 2] = num / (den + DBL_MIN);
}

// Per view: w -= alpha g, partial sums of |alpha g| and |w_new|
// (steepest_descent_iter, :228-237).
__global__ __launch_bounds__(kThreads) void k_pr_update(PrArgs P, double* __restrict__ part) {
  const int view = blockIdx.x / kPrParts, blk = blockIdx.x % kPrParts;
  if (!P.active[view]) return;
  const long long o = (long long)view * P.npix;
  long long j0, j1;
  pr_chunk(P.npix, blk, j0, j1);
  const float alpha = (float)P.scal[4 * view + 2];
  double up = 0.0, val = 0.0;
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const float2 g = P.g[o + j];
    const float2 w = csub(P.w[o + j], cscale(g, alpha));
    P.w[o + j] = w;
    up += fabs((double)alpha) * sqrt((double)g.x * g.x + (double)g.y * g.y);
    val += sqrt((double)w.x * w.x + (double)w.y * w.y);
  }
  block_sum2(up, val);
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = up;
    part[2 * (size_t)blockIdx.x + 1] = val;
  }
}

// upd = 100 sum|alpha g| / sum|w_new| per view from the partials.
__global__ void k_pr_update_fin(PrArgs P, const double* __restrict__ part, int n_views) {
  const int view = blockIdx.x * blockDim.x + threadIdx.x;
  if (view >= n_views || !P.active[view]) return;
  double up = 0.0, val = 0.0;
  for (int b = 0; b < kPrParts; ++b) {
    up += part[2 * ((size_t)view * kPrParts + b)];
    val += part[2 * ((size_t)view * kPrParts + b) + 1];
  }
  P.scal[4 * view + 3] = 100.0 * up / val;
}

// Omega = y conj(h) / (|h|^2 + eps), normalised (compute_phase_projection,
// :10-33); h = H D w in b1.
__global__ void k_pr_omega(PrArgs P) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.nall;
       i += (long long)gridDim.x * blockDim.x) {
    if (!P.active[i / P.npix]) continue;
    const float2 h = P.b1[i];
    const float2 q = cscale(cmulc(P.y[i], h), 1.f / (h.x * h.x + h.y * h.y + FLT_MIN));
    P.om[i] = cscale(q, 1.f / (sqrtf(q.x * q.x + q.y * q.y) + FLT_MIN));
  }
}

// Per view cost: |y - Omega h|^2 Lambda / 2 + nu |w - v|^2 / 2, h = H D w in
// b1 (compute_cost, :172-205).
__global__ __launch_bounds__(kThreads) void k_pr_cost(PrArgs P, double* __restrict__ out) {
  const int view = blockIdx.x;
  const long long o = (long long)view * P.npix;
  double c1 = 0.0, c2 = 0.0;
  for (long long j = threadIdx.x; j < P.npix; j += kThreads) {
    const float2 r = csub(P.y[o + j], cmul(P.b1[o + j], P.om[o + j]));
    const float2 d = csub(P.w[o + j], P.v[o + j]);
    c1 += ((double)r.x * r.x + (double)r.y * r.y) * P.lam[o + j];
    c2 += (double)d.x * d.x + (double)d.y * d.y;
  }
  block_sum2(c1, c2);
  if (threadIdx.x == 0) out[view] = c1 / 2 + (double)P.nu * c2 / 2;
}

// Stopping test of XT_CmplxProjEst.c:165-169 after iteration j.
__global__ void k_pr_stop(int n_views, const double* __restrict__ scal, int* __restrict__ active,
                          int* __restrict__ iters, double thresh, int j) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_views || !active[v]) return;
  iters[v] = j + 1;
  if (scal[4 * v + 3] < thresh && j > 1) active[v] = 0;
}

// ---------------------------------------------------------------- NMS
// Parameters of the projection-domain prior of Cost_NMS (XT_Search.c:8-43,
// without ENABLE_TOMO_RECONS, Q12) and of the simplex search; tomo selects
// the ENABLE_TOMO_RECONS cost, mu/2 |z - a|^2 towards the tomographic
// projection a in place of the prior (XT_Search.c:18-19).
struct NmsParams {
  double T[2][2];     // decorrelating transform: row 0 phase, row 1 magnitude
  double filt[9];     // in-plane 3x3 neighbour weights
  double c[2], sq[2], sqp[2];  // per prior (0 phase, 1 magnitude): c, sigma^2, sigma^(2-p)
  double sqc[2], sqr[2];       // c sigma^2 and sigma^2 / sigma^(2-p)
  double qmp;         // q - p
  double nu;
  double rho, chi, gamma, sigma, thresh;
  int max_iter;
  int tomo;
  double mu;
};

// One pixel's data: b = w - v and its 3x3 neighbourhood (flags bit k set
// where neighbour k = 3 m + n lies inside the view).
struct NmsPix {
  double nr[9], ni[9];
  double br, bi;
  double ar, ai;  // tomographic projection (tomo only)
  unsigned flags;
};

// qGGMRF value with q = 2 (Mag_/Phase_QGGMRF_Spatial_Value, XT_Prior.c:7-15):
// (|d|^2 / s^q) / (c + |d|^(q-p) / s^(q-p)), written with one division as
// |d|^2 / (c s^q + |d|^(q-p) s^q / s^(q-p)) (sqc, sqr set on the host), and
// |d|^(q-p) as exp2((q - p) log2 |d|): two of the device library's FP64
// functions in place of its general pow.  Within a few ulp of the
// reference's form; the search paths of the parity test are unchanged.
__device__ __forceinline__ double nms_value(double d, const NmsParams& P, int t) {
  const double ad = fabs(d);
  const double r = ad == 0.0 ? 0.0 : exp2(P.qmp * log2(ad));
  return (ad * ad) / (P.sqc[t] + r * P.sqr[t]);
}

// Cost_NMS (XT_Search.c:8-43), FP64, terms in the reference's order.
__device__ __forceinline__ double nms_cost(double zr, double zi, const NmsPix& X,
                                          const NmsParams& P) {
  const double e = exp(-zr);
  double sn, cs;
  sincos(-zi, &sn, &cs);
  const double cv = e * cs, sv = e * sn;
  double cost = ((cv - X.br) * (cv - X.br) + (sv - X.bi) * (sv - X.bi)) * P.nu / 2;
  if (P.tomo) return cost + ((zr - X.ar) * (zr - X.ar) + (zi - X.ai) * (zi - X.ai)) * P.mu / 2;
  // unrolled, so that the neighbourhood stays in registers
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (!((X.flags >> k) & 1u)) continue;
    const double db = zr - X.nr[k], dd = zi - X.ni[k];
    cost += P.filt[k] * nms_value(P.T[0][0] * dd + P.T[0][1] * db, P, 0);
    cost += P.filt[k] * nms_value(P.T[1][0] * dd + P.T[1][1] * db, P, 1);
  }
  return cost;
}

struct Simplex {
  double r[4], i[4], c[4];  // [0] centroid, [1] best, [2] second, [3] worst
  __device__ void swap(int a, int b) {
    double t = r[a]; r[a] = r[b]; r[b] = t;
    t = i[a]; i[a] = i[b]; i[b] = t;
    t = c[a]; c[a] = c[b]; c[b] = t;
  }
  __device__ void sort() {  // NMS_Sort
    if (c[3] < c[2]) swap(2, 3);
    if (c[2] < c[1]) swap(1, 2);
    if (c[3] < c[2]) swap(2, 3);
  }
  __device__ void accept(double zr, double zi, double cost) {  // NMS_Accept
    r[3] = zr; i[3] = zi; c[3] = cost;
    if (c[3] < c[2]) swap(2, 3);
    if (c[2] < c[1]) swap(1, 2);
  }
};

// Nelder_Mead_Simplex_2DSearch (XT_Search.c:100-215) from (zr, zi); returns
// the iterations run and leaves the best vertex in (zr, zi).  Each iteration
// evaluates the reflection, then at most one more point -- the expansion
// (reflection below the best), the outside contraction (between the second
// and the worst) or the inside contraction (at or above the worst) -- chosen
// per pixel before a single cost evaluation, and the shrink's two points:
// the reference's decisions and arithmetic in an order where the lanes of a
// wavefront that took different branches share one evaluation.
__device__ int nms_search(double& zr, double& zi, const NmsPix& X, const NmsParams& P) {
  Simplex S;
  S.r[1] = zr; S.i[1] = zi;
  S.r[2] = 2 * zr; S.i[2] = zi;
  S.r[3] = zr; S.i[3] = 2 * zi;
#pragma unroll 1
  for (int k = 1; k < 4; ++k) {
    const double c = nms_cost(k == 2 ? 2 * zr : zr, k == 3 ? 2 * zi : zi, X, P);
    if (k == 1) S.c[1] = c;
    else if (k == 2) S.c[2] = c;
    else S.c[3] = c;
  }
  S.sort();
  S.r[0] = (S.r[1] + S.r[2]) / 2.0;
  S.i[0] = (S.i[1] + S.i[2]) / 2.0;
  int it = 0;
  for (; it < P.max_iter; ++it) {
    double ocr[4], oci[4];
    for (int k = 0; k < 4; ++k) { ocr[k] = S.r[k]; oci[k] = S.i[k]; }
    const double rr = S.r[0] + P.rho * (S.r[0] - S.r[3]);
    const double ri = S.i[0] + P.rho * (S.i[0] - S.i[3]);
    const double cr = nms_cost(rr, ri, X, P);
    // -1 none (incomparable costs), 0 accept the reflection, 1 expand,
    // 2 outside contraction, 3 inside contraction
    int kind = -1;
    if (cr < S.c[2] && cr >= S.c[1]) kind = 0;
    else if (cr < S.c[1]) kind = 1;
    else if (cr >= S.c[2] && cr < S.c[3]) kind = 2;
    else if (cr >= S.c[3]) kind = 3;
    double pr = rr, pi = ri;
    if (kind == 1) {
      pr = S.r[0] + P.chi * (rr - S.r[0]);
      pi = S.i[0] + P.chi * (ri - S.i[0]);
    } else if (kind == 2) {
      pr = S.r[0] + P.gamma * (rr - S.r[0]);
      pi = S.i[0] + P.gamma * (ri - S.i[0]);
    } else if (kind == 3) {
      pr = S.r[0] - P.gamma * (S.r[0] - S.r[3]);
      pi = S.i[0] - P.gamma * (S.i[0] - S.i[3]);
    }
    double cp = 0.0;
    if (kind >= 1) cp = nms_cost(pr, pi, X, P);
    bool shrink = false;
    if (kind == 0) {
      S.accept(rr, ri, cr);
    } else if (kind == 1) {
      if (cp < cr) S.accept(pr, pi, cp);
      else S.accept(rr, ri, cr);
    } else if (kind == 2) {
      if (cp <= cr) S.accept(pr, pi, cp);
      else shrink = true;
    } else if (kind == 3) {
      if (cp < S.c[3]) S.accept(pr, pi, cp);
      else shrink = true;
    }
    if (shrink) {  // NMS_Shrink
      S.r[2] = S.r[1] + P.sigma * (S.r[2] - S.r[1]);
      S.i[2] = S.i[1] + P.sigma * (S.i[2] - S.i[1]);
      S.r[3] = S.r[1] + P.sigma * (S.r[3] - S.r[1]);
      S.i[3] = S.i[1] + P.sigma * (S.i[3] - S.i[1]);
#pragma unroll 1
      for (int k = 2; k < 4; ++k) {
        const double c = nms_cost(k == 2 ? S.r[2] : S.r[3], k == 2 ? S.i[2] : S.i[3], X, P);
        if (k == 2) S.c[2] = c;
        else S.c[3] = c;
      }
      S.sort();
    }
    S.r[0] = (S.r[1] + S.r[2]) / 2.0;
    S.i[0] = (S.i[1] + S.i[2]) / 2.0;
    double num = 0.0, den = 0.0;
    for (int k = 0; k < 4; ++k) {
      num += fabs(ocr[k] - S.r[k]) + fabs(oci[k] - S.i[k]);
      den += fabs(ocr[k]) + fabs(oci[k]);
    }
    if (num / den * 100 < P.thresh && it > 1) break;
  }
  zr = S.r[1];
  zi = S.i[1];
  return it;
}

// One colour (j%2, k%2) = (cj, ck) of the four-colour NMS pass: every pixel
// of that colour of every active view, one thread each; pixels of a colour
// are not neighbours, so they update independently.  b = w - v from w, vd.
__global__ __launch_bounds__(kThreads) void k_nms_colour(
    NmsParams P, double* __restrict__ zr, double* __restrict__ zi, const float2* __restrict__ w,
    const float2* __restrict__ vd, const int* __restrict__ active, const double* __restrict__ ar,
    const double* __restrict__ ai, int n_views, int rows, int cols, int cj, int ck,
    int* __restrict__ iters) {
  const int hr = (rows - cj + 1) / 2, hc = (cols - ck + 1) / 2;
  const long long per = (long long)hr * hc;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * n_views) return;
  const int view = (int)(t / per);
  if (!active[view]) return;
  const long long q = t % per;
  const int j = cj + 2 * (int)(q / hc), k = ck + 2 * (int)(q % hc);
  const long long o = (long long)view * rows * cols;
  NmsPix X;
  X.flags = 0;
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) {
      const int a = j + m - 1, b = k + n - 1, kk = 3 * m + n;
      const bool in = a >= 0 && a < rows && b >= 0 && b < cols;
      X.nr[kk] = in ? zr[o + (long long)a * cols + b] : 0.0;
      X.ni[kk] = in ? zi[o + (long long)a * cols + b] : 0.0;
      if (in) X.flags |= 1u << kk;
    }
  const long long p = o + (long long)j * cols + k;
  X.br = (double)w[p].x - (double)vd[p].x;
  X.bi = (double)w[p].y - (double)vd[p].y;
  X.ar = P.tomo ? ar[p] : 0.0;
  X.ai = P.tomo ? ai[p] : 0.0;
  double r = zr[p], i = zi[p];
  const int it = nms_search(r, i, X, P);
  zr[p] = r;
  zi[p] = i;
  iters[p] = it;
}

// exp(-z) (+ add): the transmittance of XT_CmplxProjEst.c:175-178.
__device__ __forceinline__ float2 transmit(double zr, double zi) {
  const double e = exp(-zr);
  return make_float2((float)(e * cos(-zi)), (float)(e * sin(-zi)));
}

// out = D exp(-z) for active views (the argument of compute_pretcost's H).
__global__ void k_dexpz(const double* __restrict__ zr, const double* __restrict__ zi,
                        const float2* __restrict__ D, const int* __restrict__ active,
                        long long npix, long long nall, float2* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nall;
       i += (long long)gridDim.x * blockDim.x) {
    if (!active[i / npix]) continue;
    out[i] = cmul(transmit(zr[i], zi[i]), D[i]);
  }
}

// SD target exp(-z) + vd and the copy w_old = w (XT_CmplxProjEst.c:171-179).
__global__ void k_sd_target(const double* __restrict__ zr, const double* __restrict__ zi,
                            const float2* __restrict__ vd, const float2* __restrict__ w,
                            const int* __restrict__ active, long long npix, long long nall,
                            float2* __restrict__ target, float2* __restrict__ wold) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nall;
       i += (long long)gridDim.x * blockDim.x) {
    if (!active[i / npix]) continue;
    target[i] = cadd(transmit(zr[i], zi[i]), vd[i]);
    wold[i] = w[i];
  }
}

// Per view: the pretcost (XT_CmplxProjEst.c:15-103) -- the forward term from
// h = H(D exp(-z)) in hb and the prior over right / down / down-right /
// down-left pairs, or with tomo mu/2 |z - a|^2 (XT_CmplxProjEst.c:38-41).
__global__ __launch_bounds__(kThreads) void k_pret_cost(NmsParams P, const double* __restrict__ zr,
                                                        const double* __restrict__ zi,
                                                        const float2* __restrict__ y,
                                                        const float* __restrict__ lam,
                                                        const float2* __restrict__ hb,
                                                        const int* __restrict__ active,
                                                        const double* __restrict__ ar,
                                                        const double* __restrict__ ai, int rows,
                                                        int cols, double* __restrict__ out) {
  const int view = blockIdx.x;
  if (!active[view]) return;
  const long long o = (long long)view * rows * cols;
  double fwd = 0.0, prior = 0.0;
  const int dj[4] = {0, 1, 1, 1}, dk[4] = {1, 0, 1, -1}, fk[4] = {5, 7, 8, 6};
  for (long long p = threadIdx.x; p < (long long)rows * cols; p += kThreads) {
    const float2 h = hb[o + p];
    const double r = (double)y[o + p].x - sqrt((double)h.x * h.x + (double)h.y * h.y);
    fwd += r * r * lam[o + p] / 2;
    if (P.tomo) {
      const double dr = zr[o + p] - ar[o + p], di = zi[o + p] - ai[o + p];
      prior += P.mu * (dr * dr + di * di) / 2;
      continue;
    }
    const int j = (int)(p / cols), k = (int)(p % cols);
    for (int e = 0; e < 4; ++e) {
      const int a = j + dj[e], b = k + dk[e];
      if (a >= rows || b < 0 || b >= cols) continue;
      const long long q = o + (long long)a * cols + b;
      const double db = zr[o + p] - zr[q], dd = zi[o + p] - zi[q];
      prior += P.filt[fk[e]] * nms_value(P.T[0][0] * dd + P.T[0][1] * db, P, 0);
      prior += P.filt[fk[e]] * nms_value(P.T[1][0] * dd + P.T[1][1] * db, P, 1);
    }
  }
  block_sum2(fwd, prior);
  if (threadIdx.x == 0) out[view] = fwd + prior;
}

// Per view: the dual update vd += exp(-z) - w and the residual sums
// (XT_CmplxProjEst.c:181-197); res[3 view + 0..2] = primal, dual, |v|^2.
__global__ __launch_bounds__(kThreads) void k_dual(const double* __restrict__ zr,
                                                   const double* __restrict__ zi,
                                                   const float2* __restrict__ w,
                                                   const float2* __restrict__ wold,
                                                   float2* __restrict__ vd,
                                                   const int* __restrict__ active, long long npix,
                                                   double* __restrict__ res) {
  const int view = blockIdx.x;
  if (!active[view]) return;
  const long long o = (long long)view * npix;
  double pr = 0.0, du = 0.0, vs = 0.0;
  for (long long p = threadIdx.x; p < npix; p += kThreads) {
    const float2 t = csub(transmit(zr[o + p], zi[o + p]), w[o + p]);
    const float2 v = cadd(vd[o + p], t);
    vd[o + p] = v;
    const float2 dw = csub(w[o + p], wold[o + p]);
    pr += (double)t.x * t.x + (double)t.y * t.y;
    du += (double)dw.x * dw.x + (double)dw.y * dw.y;
    vs += (double)v.x * v.x + (double)v.y * v.y;
  }
  block_sum2(pr, du);
  double vs0 = vs, dummy = 0.0;
  block_sum2(vs0, dummy);
  if (threadIdx.x == 0) {
    res[3 * view + 0] = sqrt(pr / (double)npix);
    res[3 * view + 1] = sqrt(du / vs0);
    res[3 * view + 2] = vs0;
  }
}

// Paganin input: the intensity ratio |y|^2 / |D|^2 as a real-valued complex
// array (paganins_1mat_phase_retrieval, XT_Paganin.c:84-93; the reference
// squares the real part of the measurement amplitude).
__global__ void k_pag_pre(const float2* __restrict__ y, const float2* __restrict__ D,
                          long long nall, float2* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nall;
       i += (long long)gridDim.x * blockDim.x) {
    const float2 d = D[i];
    const double i0 = (double)d.x * d.x + (double)d.y * d.y;
    out[i] = make_float2((float)((double)y[i].x * y[i].x / i0), 0.f);
  }
}

// Paganin output: thickness -log|Re| of the filtered intensity, z_real =
// thick / 2, z_imag = thick delta/beta / 2 (XT_Paganin.c:117-131).
__global__ void k_pag_post(const float2* __restrict__ f, long long nall, double dob,
                           double* __restrict__ zr, double* __restrict__ zi) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nall;
       i += (long long)gridDim.x * blockDim.x) {
    const double thick = -log(fabs((double)f[i].x));
    zr[i] = thick / 2;
    zi[i] = thick * dob / 2;
  }
}

__global__ void k_fill_int(int* p, int n, int val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = val;
}

template <typename T>
int dev_alloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIP_OK(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return (1 << l) == v ? l : -1;
}

}  // namespace

struct pcmbir_hip_ctx {
  int nv = 0, rows = 0, cols = 0, logr = 0, logc = 0;
  long long npix = 0, nall = 0;
  hipStream_t stream = nullptr;
  float2 *tw_r = nullptr, *tw_c = nullptr, *fac_h = nullptr, *fac_hh = nullptr;
  float2* fac_pag = nullptr;  // Paganin filter (built on first use)
  double pag_dob = -1.0, pag_k = 0.0;
  double delta_rows = 0.0, delta_cols = 0.0, wavelength = 0.0, dist = 0.0;
  float2 *y = nullptr, *D = nullptr, *v = nullptr, *w = nullptr, *om = nullptr, *g = nullptr;
  float2* b = nullptr;  // b1 | b2, contiguous: one batched transform of 2 n_views
  float* lam = nullptr;
  int *active = nullptr, *iters = nullptr;
  double *scal = nullptr, *cost = nullptr;
  double* part = nullptr;  // [n_views][kPrParts][2] partial sums of the step's reductions
  float nu = 0.f;
  bool problem = false;
  // NMS / ADMM state (estimate_complex_projection)
  NmsParams nms;
  bool nms_set = false;
  double *zr = nullptr, *zi = nullptr, *res = nullptr;
  float2 *vd = nullptr, *wold = nullptr;
  int *pix_iters = nullptr, *admm = nullptr;
  // tomographic projection a of the ENABLE_TOMO_RECONS coupling
  double *ar = nullptr, *ai = nullptr;
  bool tomo = false;
  double tomo_mu = 0.0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t ev_used = 0;
  double fft_ms = 0.0;
  int64_t fft_launches = 0;
  bool fft_reg = true;  // register-resident FFT passes (option; else the LDS radix-2 ones)
  int col_threads = 512;  // most threads per workgroup of the register column pass
};

namespace {

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n / 2 > 0 ? n / 2 : 1);
  for (int m = 0; m < n / 2; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)n;
    t[m] = make_float2((float)cos(a), (float)sin(a));
  }
  return t;
}

// FFT, multiplication by the spectral factor fac [rows][cols], inverse FFT
// (unnormalised; fac carries 1/N) of nbatch views stored at buf.
int filter_dev(pcmbir_hip_ctx* c, float2* buf, int nbatch, const float2* fac) {
  if (c->ev_used == c->ev.size()) {
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    c->ev.push_back({a, b});
  }
  HIP_OK(hipEventRecord(c->ev[c->ev_used].first, c->stream));
  const long long nrows = (long long)nbatch * c->rows;
  const unsigned rb = (unsigned)((nrows + (kRowElems >> c->logc) - 1) / (kRowElems >> c->logc));
  // register-resident passes where a sequence holds at least 8 values
  const bool rows_reg = c->fft_reg && c->logc >= 3, cols_reg = c->fft_reg && c->logr >= 3;
  if (rows_reg)
    hipLaunchKernelGGL(k_fft_rows_reg<false>, dim3(rb), dim3(256), 0, c->stream, buf, nrows,
                       c->logc, c->tw_c);
  else
    hipLaunchKernelGGL(k_fft_rows, dim3(rb), dim3(kThreads), 0, c->stream, buf, nrows, c->logc,
                       c->tw_c, 0);
  if (cols_reg) {
    const int ltc = col_tile_reg(c->logr, c->logc, c->col_threads);
    const unsigned threads = 1u << (ltc + c->logr - 3);
    const size_t lds = (sizeof(float2) << (ltc + c->logr)) + (sizeof(float2) << (c->logr - 1));
    hipLaunchKernelGGL(k_fresnel_cols_reg, dim3((unsigned)nbatch << (c->logc - ltc)),
                       dim3(threads), lds, c->stream, buf, c->logr, c->logc, ltc, c->tw_r, fac);
  } else {
    const int tiles = c->cols / col_tile(c->logr, c->logc);
    hipLaunchKernelGGL(k_fresnel_cols, dim3((unsigned)(nbatch * tiles)), dim3(kThreads), 0,
                       c->stream, buf, c->logr, c->logc, c->tw_r, fac);
  }
  if (rows_reg)
    hipLaunchKernelGGL(k_fft_rows_reg<true>, dim3(rb), dim3(256), 0, c->stream, buf, nrows,
                       c->logc, c->tw_c);
  else
    hipLaunchKernelGGL(k_fft_rows, dim3(rb), dim3(kThreads), 0, c->stream, buf, nrows, c->logc,
                       c->tw_c, 1);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(c->ev[c->ev_used].second, c->stream));
  c->ev_used++;
  c->fft_launches += 3;
  return 0;
}

// Fresnel transform (herm = 0: H, 1: H^H) of nbatch views stored at buf.
int fresnel_dev(pcmbir_hip_ctx* c, float2* buf, int nbatch, int herm) {
  return filter_dev(c, buf, nbatch, herm ? c->fac_hh : c->fac_h);
}

int collect_events(pcmbir_hip_ctx* c) {
  HIP_OK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < c->ev_used; ++i) {
    float m = 0.f;
    HIP_OK(hipEventElapsedTime(&m, c->ev[i].first, c->ev[i].second));
    c->fft_ms += m;
  }
  c->ev_used = 0;
  return 0;
}

unsigned ew_blocks(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  return (unsigned)(b > 4096 ? 4096 : b);
}

PrArgs pr_args(pcmbir_hip_ctx* c) {
  PrArgs P;
  P.y = c->y;
  P.D = c->D;
  P.v = c->v;
  P.lam = c->lam;
  P.w = c->w;
  P.om = c->om;
  P.g = c->g;
  P.b1 = c->b;
  P.b2 = c->b + c->nall;
  P.active = c->active;
  P.scal = c->scal;
  P.nu = c->nu;
  P.npix = c->npix;
  P.nall = c->nall;
  return P;
}

// Omega = phase projection of the current w (compute_phase_projection,
// XT_PhaseRet.c:10-33), every view.
int omega_from_w(pcmbir_hip_ctx* c) {
  const unsigned eb = ew_blocks(c->nall), vb = (unsigned)((c->nv + 255) / 256);
  PrArgs P = pr_args(c);
  hipLaunchKernelGGL(k_fill_int, dim3(vb), dim3(256), 0, c->stream, c->active, c->nv, 1);
  hipLaunchKernelGGL(k_pr_dmul, dim3(eb), dim3(kThreads), 0, c->stream, P, c->w, P.b1);
  if (fresnel_dev(c, P.b1, c->nv, 0)) return -1;
  hipLaunchKernelGGL(k_pr_omega, dim3(eb), dim3(kThreads), 0, c->stream, P);
  HIP_OK(hipGetLastError());
  return 0;
}

// The steepest-descent loop over every view whose init_active flag is set
// (all views when init_active is NULL); per-view stopping as
// XT_CmplxProjEst.c:165-169.
int sd_loop(pcmbir_hip_ctx* c, int max_iters, double thresh, const int* init_active) {
  const PrArgs P = pr_args(c);
  const unsigned eb = ew_blocks(c->nall);
  const unsigned vb = (unsigned)((c->nv + 255) / 256);
  if (init_active)
    HIP_OK(hipMemcpyAsync(c->active, init_active, c->nv * sizeof(int), hipMemcpyDeviceToDevice,
                          c->stream));
  else
    hipLaunchKernelGGL(k_fill_int, dim3(vb), dim3(256), 0, c->stream, c->active, c->nv, 1);
  hipLaunchKernelGGL(k_fill_int, dim3(vb), dim3(256), 0, c->stream, c->iters, c->nv, 0);
  HIP_OK(hipGetLastError());
  std::vector<int> act(c->nv);
  for (int j = 0; j < max_iters; ++j) {
    // gradient (compute_gradient): b1 = H D w, then the two Hermitian passes
    hipLaunchKernelGGL(k_pr_dmul, dim3(eb), dim3(kThreads), 0, c->stream, P, c->w, P.b1);
    if (fresnel_dev(c, P.b1, c->nv, 0)) return -1;
    hipLaunchKernelGGL(k_pr_mid, dim3(eb), dim3(kThreads), 0, c->stream, P, 1);
    if (fresnel_dev(c, P.b1, 2 * c->nv, 1)) return -1;
    hipLaunchKernelGGL(k_pr_grad, dim3(eb), dim3(kThreads), 0, c->stream, P);
    // step size (compute_stepsize)
    hipLaunchKernelGGL(k_pr_dmul, dim3(eb), dim3(kThreads), 0, c->stream, P, c->g, P.b1);
    if (fresnel_dev(c, P.b1, c->nv, 0)) return -1;
    hipLaunchKernelGGL(k_pr_mid, dim3(eb), dim3(kThreads), 0, c->stream, P, 0);
    if (fresnel_dev(c, P.b1, c->nv, 1)) return -1;
    hipLaunchKernelGGL(k_pr_step, dim3(c->nv * kPrParts), dim3(kThreads), 0, c->stream, P,
                       c->part);
    hipLaunchKernelGGL(k_pr_step_fin, dim3(vb), dim3(256), 0, c->stream, P, c->part, c->nv);
    hipLaunchKernelGGL(k_pr_update, dim3(c->nv * kPrParts), dim3(kThreads), 0, c->stream, P,
                       c->part);
    hipLaunchKernelGGL(k_pr_update_fin, dim3(vb), dim3(256), 0, c->stream, P, c->part, c->nv);
    // phase projection
    hipLaunchKernelGGL(k_pr_dmul, dim3(eb), dim3(kThreads), 0, c->stream, P, c->w, P.b1);
    if (fresnel_dev(c, P.b1, c->nv, 0)) return -1;
    hipLaunchKernelGGL(k_pr_omega, dim3(eb), dim3(kThreads), 0, c->stream, P);
    hipLaunchKernelGGL(k_pr_stop, dim3(vb), dim3(256), 0, c->stream, c->nv, c->scal, c->active,
                       c->iters, thresh, j);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(act.data(), c->active, c->nv * sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    if (collect_events(c)) return -1;
    bool any = false;
    for (int a : act) any = any || a;
    if (!any) break;
  }
  return 0;
}

int alloc_nms_state(pcmbir_hip_ctx* c) {
  if (c->zr) return 0;
  if (dev_alloc(&c->zr, c->nall) || dev_alloc(&c->zi, c->nall) || dev_alloc(&c->res, 3 * (size_t)c->nv) ||
      dev_alloc(&c->vd, c->nall) || dev_alloc(&c->wold, c->nall) ||
      dev_alloc(&c->pix_iters, c->nall) || dev_alloc(&c->admm, c->nv))
    return -1;
  return 0;
}

// The four colours of an NMS pass over the active views (admm mask).
int nms_pass_dev(pcmbir_hip_ctx* c, const float2* w, const float2* vd, const int* mask) {
  for (int col = 0; col < 4; ++col) {
    const int cj = col >> 1, ck = col & 1;
    const long long per =
        (long long)((c->rows - cj + 1) / 2) * ((c->cols - ck + 1) / 2) * c->nv;
    hipLaunchKernelGGL(k_nms_colour, dim3((unsigned)((per + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, c->stream, c->nms, c->zr, c->zi, w, vd, mask, c->ar,
                       c->ai, c->nv, c->rows, c->cols, cj, ck, c->pix_iters);
  }
  HIP_OK(hipGetLastError());
  return 0;
}

// Per-view pretcost of the current z into out (active views only).
int pret_cost_dev(pcmbir_hip_ctx* c, const int* mask, double* out) {
  const unsigned eb = ew_blocks(c->nall);
  hipLaunchKernelGGL(k_dexpz, dim3(eb), dim3(kThreads), 0, c->stream, c->zr, c->zi, c->D, mask,
                     c->npix, c->nall, c->b);
  if (fresnel_dev(c, c->b, c->nv, 0)) return -1;
  hipLaunchKernelGGL(k_pret_cost, dim3(c->nv), dim3(kThreads), 0, c->stream, c->nms, c->zr, c->zi,
                     c->y, c->lam, c->b, mask, c->ar, c->ai, c->rows, c->cols, out);
  HIP_OK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

const char* pcmbir_hip_last_error(void) { return g_err.c_str(); }

int pcmbir_hip_create(pcmbir_hip_ctx** out, int32_t n_views, int32_t rows, int32_t cols,
                      double delta_rows, double delta_cols, double lam, double dist,
                      const float* window) {
  if (!out) return fail("null output pointer");
  *out = nullptr;
  const int lr = ilog2(rows), lc = ilog2(cols);
  if (n_views <= 0 || lr < 1 || lc < 1 || rows > kFftMax || cols > kFftMax)
    return fail("rows and cols must be powers of two in [2, 1024], n_views > 0");
  if (!(delta_rows > 0) || !(delta_cols > 0)) return fail("sample spacings must be positive");
  auto* c = new pcmbir_hip_ctx();
  c->nv = n_views;
  c->rows = rows;
  c->cols = cols;
  c->logr = lr;
  c->logc = lc;
  c->npix = (long long)rows * cols;
  c->nall = c->npix * n_views;
  c->delta_rows = delta_rows;
  c->delta_cols = delta_cols;
  c->wavelength = lam;
  c->dist = dist;
  // spectral factors in FP64 (XT_FresnelTran.c:33-50 and :104-118), with 1/N
  const double N = (double)rows * cols;
  std::vector<float2> fh(c->npix), fhh(c->npix);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      const double u = (i >= rows / 2 ? -(double)(rows - i) : (double)i) / (rows * delta_rows);
      const double v = (j >= cols / 2 ? -(double)(cols - j) : (double)j) / (cols * delta_cols);
      const double phase = -(u * u + v * v) * M_PI * lam * dist;
      const double wv = window ? (double)window[(size_t)i * cols + j] : 1.0;
      fh[(size_t)i * cols + j] =
          make_float2((float)(cos(phase) * wv / N), (float)(sin(phase) * wv / N));
      fhh[(size_t)i * cols + j] = make_float2((float)(cos(phase) / N), (float)(-sin(phase) / N));
    }
  const std::vector<float2> tr = twiddles(rows), tc = twiddles(cols);
  int rc = 0;
  if (hipStreamCreate(&c->stream) != hipSuccess) rc = fail("hipStreamCreate failed");
  rc = rc || dev_alloc(&c->tw_r, tr.size()) || dev_alloc(&c->tw_c, tc.size()) ||
       dev_alloc(&c->fac_h, c->npix) || dev_alloc(&c->fac_hh, c->npix) ||
       dev_alloc(&c->b, 2 * c->nall);
  if (!rc) {
    rc = hipMemcpy(c->tw_r, tr.data(), tr.size() * sizeof(float2), hipMemcpyHostToDevice) ||
         hipMemcpy(c->tw_c, tc.data(), tc.size() * sizeof(float2), hipMemcpyHostToDevice) ||
         hipMemcpy(c->fac_h, fh.data(), fh.size() * sizeof(float2), hipMemcpyHostToDevice) ||
         hipMemcpy(c->fac_hh, fhh.data(), fhh.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (rc) fail("upload of the Fresnel tables failed");
  }
  if (rc) {
    std::string keep = g_err;
    pcmbir_hip_destroy(c);
    return fail(keep);
  }
  *out = c;
  return 0;
}

void pcmbir_hip_destroy(pcmbir_hip_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->tw_r, c->tw_c, c->fac_h, c->fac_hh, c->fac_pag, c->y, c->D, c->v, c->w, c->om,
                  c->g, c->b, c->lam, c->active, c->iters, c->scal, c->cost, c->part,
                  c->zr, c->zi, c->res, c->vd, c->wold, c->pix_iters, c->admm, c->ar, c->ai};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto& pr : c->ev) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int pcmbir_hip_fresnel(pcmbir_hip_ctx* c, int32_t herm, const float* x, float* y) {
  if (!c || !x || !y) return fail("null argument");
  const size_t bytes = (size_t)c->nall * sizeof(float2);
  HIP_OK(hipMemcpyAsync(c->b, x, bytes, hipMemcpyHostToDevice, c->stream));
  if (fresnel_dev(c, c->b, c->nv, herm != 0)) return -1;
  HIP_OK(hipMemcpyAsync(y, c->b, bytes, hipMemcpyDeviceToHost, c->stream));
  return collect_events(c);
}

int pcmbir_hip_set_phase_problem(pcmbir_hip_ctx* c, const float* y, const float* D,
                                 const float* lambda, const float* v, const float* w,
                                 const float* omega, double nu) {
  if (!c || !y || !D || !lambda || !v || !w) return fail("null argument");
  if (!c->y) {
    if (dev_alloc(&c->y, c->nall) || dev_alloc(&c->D, c->nall) || dev_alloc(&c->v, c->nall) ||
        dev_alloc(&c->w, c->nall) || dev_alloc(&c->om, c->nall) || dev_alloc(&c->g, c->nall) ||
        dev_alloc(&c->lam, c->nall) || dev_alloc(&c->active, c->nv) ||
        dev_alloc(&c->iters, c->nv) || dev_alloc(&c->scal, 4 * (size_t)c->nv) ||
        dev_alloc(&c->cost, c->nv) || dev_alloc(&c->part, 2 * (size_t)kPrParts * c->nv))
      return -1;
  }
  const size_t cb = (size_t)c->nall * sizeof(float2);
  HIP_OK(hipMemcpyAsync(c->y, y, cb, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->D, D, cb, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->v, v, cb, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->w, w, cb, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->lam, lambda, (size_t)c->nall * sizeof(float), hipMemcpyHostToDevice,
                        c->stream));
  if (omega) {
    HIP_OK(hipMemcpyAsync(c->om, omega, cb, hipMemcpyHostToDevice, c->stream));
  } else if (omega_from_w(c)) {
    return -1;
  }
  c->nu = (float)nu;
  c->problem = true;
  return collect_events(c);
}

int pcmbir_hip_sd_run(pcmbir_hip_ctx* c, int32_t max_iters, double thresh, int32_t* iters,
                      double* upd) {
  if (!c) return fail("null context");
  if (!c->problem) return fail("no phase problem set");
  if (max_iters < 0) return fail("max_iters must be >= 0");
  if (sd_loop(c, max_iters, thresh, nullptr)) return -1;
  if (iters) HIP_OK(hipMemcpyAsync(iters, c->iters, c->nv * sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
  if (upd) {
    std::vector<double> s(4 * (size_t)c->nv);
    HIP_OK(hipMemcpyAsync(s.data(), c->scal, s.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int v = 0; v < c->nv; ++v) upd[v] = s[4 * (size_t)v + 3];
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int pcmbir_hip_get_phase_state(pcmbir_hip_ctx* c, float* w, float* omega) {
  if (!c || !w || !omega) return fail("null argument");
  if (!c->problem) return fail("no phase problem set");
  const size_t cb = (size_t)c->nall * sizeof(float2);
  HIP_OK(hipMemcpyAsync(w, c->w, cb, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(omega, c->om, cb, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int pcmbir_hip_phase_cost(pcmbir_hip_ctx* c, double* cost) {
  if (!c || !cost) return fail("null argument");
  if (!c->problem) return fail("no phase problem set");
  PrArgs P = pr_args(c);
  const unsigned eb = ew_blocks(c->nall);
  const unsigned vb = (unsigned)((c->nv + 255) / 256);
  hipLaunchKernelGGL(k_fill_int, dim3(vb), dim3(256), 0, c->stream, c->active, c->nv, 1);
  hipLaunchKernelGGL(k_pr_dmul, dim3(eb), dim3(kThreads), 0, c->stream, P, c->w, P.b1);
  if (fresnel_dev(c, P.b1, c->nv, 0)) return -1;
  hipLaunchKernelGGL(k_pr_cost, dim3(c->nv), dim3(kThreads), 0, c->stream, P, c->cost);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(cost, c->cost, c->nv * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  return collect_events(c);
}

int pcmbir_hip_set_nms_prior(pcmbir_hip_ctx* c, double phase_sig, double phase_c,
                             double mag_sig, double mag_c, double delta_over_beta,
                             int32_t decorrelate, double p, double nu) {
  if (!c) return fail("null context");
  if (!(phase_sig > 0) || !(mag_sig > 0)) return fail("sigmas must be positive");
  NmsParams& P = c->nms;
  memset(&P, 0, sizeof(P));
  P.T[0][0] = 1.0;
  P.T[0][1] = decorrelate ? delta_over_beta : 0.0;
  P.T[1][0] = decorrelate ? 1.0 : 0.0;
  P.T[1][1] = decorrelate ? -delta_over_beta : 1.0;
  // in-plane slice of the 26-neighbour 1/distance filter (XT_Init.c:145-177)
  const double sum = 6.0 + 12.0 / sqrt(2.0) + 8.0 / sqrt(3.0);
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) {
      const double d = sqrt((m - 1.0) * (m - 1.0) + (n - 1.0) * (n - 1.0));
      P.filt[3 * m + n] = d > 0 ? 1.0 / d / sum : 0.0;
    }
  const double q = 2.0;
  P.c[0] = phase_c;
  P.c[1] = mag_c;
  P.sq[0] = pow(phase_sig, q);
  P.sq[1] = pow(mag_sig, q);
  P.sqp[0] = pow(phase_sig, q - p);
  P.sqp[1] = pow(mag_sig, q - p);
  for (int t = 0; t < 2; ++t) {
    P.sqc[t] = P.c[t] * P.sq[t];
    P.sqr[t] = P.sq[t] / P.sqp[t];
  }
  P.qmp = q - p;
  P.nu = nu;
  // simplex search constants (XT_Init.c:507-519)
  P.rho = 1.0;
  P.chi = 2.0;
  P.gamma = 0.5;
  P.sigma = 0.5;
  P.thresh = 0.005;
  P.max_iter = 50;
  P.tomo = c->tomo ? 1 : 0;
  P.mu = c->tomo_mu;
  c->nms_set = true;
  return 0;
}

int pcmbir_hip_set_tomo_target(pcmbir_hip_ctx* c, double mu, const double* a_mag,
                               const double* a_phase) {
  if (!c) return fail("null context");
  if (!a_mag || !a_phase) {
    c->tomo = false;
    c->nms.tomo = 0;
    return 0;
  }
  if (!(mu > 0)) return fail("mu must be positive");
  if (!c->ar && (dev_alloc(&c->ar, c->nall) || dev_alloc(&c->ai, c->nall))) return -1;
  const size_t db = (size_t)c->nall * sizeof(double);
  HIP_OK(hipMemcpyAsync(c->ar, a_mag, db, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->ai, a_phase, db, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->tomo = true;
  c->tomo_mu = mu;
  c->nms.tomo = 1;
  c->nms.mu = mu;
  return 0;
}

int pcmbir_hip_nms_pass(pcmbir_hip_ctx* c, double* zr, double* zi, const float* b,
                        int64_t* iters_total) {
  if (!c || !zr || !zi || !b) return fail("null argument");
  if (!c->nms_set) return fail("NMS prior not set");
  if (alloc_nms_state(c)) return -1;
  const size_t db = (size_t)c->nall * sizeof(double), cb = (size_t)c->nall * sizeof(float2);
  HIP_OK(hipMemcpyAsync(c->zr, zr, db, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->zi, zi, db, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->wold, b, cb, hipMemcpyHostToDevice, c->stream));  // b as w
  HIP_OK(hipMemsetAsync(c->vd, 0, cb, c->stream));                          // and v = 0
  const unsigned vb = (unsigned)((c->nv + 255) / 256);
  hipLaunchKernelGGL(k_fill_int, dim3(vb), dim3(256), 0, c->stream, c->admm, c->nv, 1);
  if (nms_pass_dev(c, c->wold, c->vd, c->admm)) return -1;
  HIP_OK(hipMemcpyAsync(zr, c->zr, db, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(zi, c->zi, db, hipMemcpyDeviceToHost, c->stream));
  std::vector<int> it(c->nall);
  HIP_OK(hipMemcpyAsync(it.data(), c->pix_iters, c->nall * sizeof(int), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (iters_total) {
    int64_t t = 0;
    for (int v : it) t += v;
    *iters_total = t;
  }
  return 0;
}

int pcmbir_hip_estimate(pcmbir_hip_ctx* c, double* zr, double* zi, int32_t pret_iters,
                        double pret_thresh, int32_t sd_iters, double sd_thresh, int32_t* iters,
                        double* cost_init, double* cost_final, double* primal, double* dual,
                        double* costs) {
  if (!c || !zr || !zi) return fail("null argument");
  if (!c->problem) return fail("no phase problem set");
  if (!c->nms_set) return fail("NMS prior not set");
  if (pret_iters < 0 || sd_iters < 0) return fail("iteration counts must be >= 0");
  if (alloc_nms_state(c)) return -1;
  const size_t db = (size_t)c->nall * sizeof(double), cb = (size_t)c->nall * sizeof(float2);
  HIP_OK(hipMemcpyAsync(c->zr, zr, db, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->zi, zi, db, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemsetAsync(c->vd, 0, cb, c->stream));  // v = 0 (XT_CmplxProjEst.c:117-118)
  const unsigned vb = (unsigned)((c->nv + 255) / 256), eb = ew_blocks(c->nall);
  hipLaunchKernelGGL(k_fill_int, dim3(vb), dim3(256), 0, c->stream, c->admm, c->nv, 1);
  if (pret_cost_dev(c, c->admm, c->cost)) return -1;
  std::vector<double> pc(c->nv), res(3 * (size_t)c->nv);
  HIP_OK(hipMemcpyAsync(pc.data(), c->cost, c->nv * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (cost_init)
    for (int v = 0; v < c->nv; ++v) cost_init[v] = pc[v];
  std::vector<int> act(c->nv, 1), its(c->nv, pret_iters);
  for (int it = 0; it < pret_iters; ++it) {
    if (nms_pass_dev(c, c->w, c->vd, c->admm)) return -1;
    // SD target exp(-z) + v into the problem's v; w_old = w
    hipLaunchKernelGGL(k_sd_target, dim3(eb), dim3(kThreads), 0, c->stream, c->zr, c->zi, c->vd,
                       c->w, c->admm, c->npix, c->nall, c->v, c->wold);
    if (sd_loop(c, sd_iters, sd_thresh, c->admm)) return -1;
    hipLaunchKernelGGL(k_dual, dim3(c->nv), dim3(kThreads), 0, c->stream, c->zr, c->zi, c->w,
                       c->wold, c->vd, c->admm, c->npix, c->res);
    if (pret_cost_dev(c, c->admm, c->cost)) return -1;
    HIP_OK(hipMemcpyAsync(res.data(), c->res, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipMemcpyAsync(pc.data(), c->cost, c->nv * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    bool any = false;
    for (int v = 0; v < c->nv; ++v) {
      if (!act[v]) continue;
      if (primal) primal[(size_t)v * pret_iters + it] = res[3 * (size_t)v];
      if (dual) dual[(size_t)v * pret_iters + it] = res[3 * (size_t)v + 1];
      if (costs) costs[(size_t)v * pret_iters + it] = pc[v];
      if (cost_final) cost_final[v] = pc[v];
      if (res[3 * (size_t)v + 1] < pret_thresh && it > 1) {
        act[v] = 0;
        its[v] = it + 1;
      }
      any = any || act[v];
    }
    HIP_OK(hipMemcpyAsync(c->admm, act.data(), c->nv * sizeof(int), hipMemcpyHostToDevice,
                          c->stream));
    if (!any) break;
  }
  if (iters)
    for (int v = 0; v < c->nv; ++v) iters[v] = its[v];
  HIP_OK(hipMemcpyAsync(zr, c->zr, db, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(zi, c->zi, db, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int pcmbir_hip_get_dual(pcmbir_hip_ctx* c, float* v) {
  if (!c || !v) return fail("null argument");
  const size_t cb = (size_t)c->nall * sizeof(float2);
  if (!c->vd) {
    memset(v, 0, cb);
    return 0;
  }
  HIP_OK(hipMemcpyAsync(v, c->vd, cb, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

int pcmbir_hip_paganin(pcmbir_hip_ctx* c, const float* y, const float* D,
                       double delta_over_beta, double wavenumber, double* zr, double* zi) {
  if (!c || !y || !D || !zr || !zi) return fail("null argument");
  if (!(wavenumber > 0)) return fail("wavenumber must be positive");
  if (!c->fac_pag || c->pag_dob != delta_over_beta || c->pag_k != wavenumber) {
    // 1 / ((z (delta/beta) / (2 k) 4 pi^2 (u^2 + v^2) + 1) N), FFT-shifted
    // frequencies (XT_Paganin.c:97-113), in FP64
    const double N = (double)c->rows * c->cols, reg = delta_over_beta / (2.0 * wavenumber);
    std::vector<float2> f(c->npix);
    for (int i = 0; i < c->rows; ++i)
      for (int j = 0; j < c->cols; ++j) {
        const double u = (i >= c->rows / 2 ? -(double)(c->rows - i) : (double)i) /
                         (c->rows * c->delta_rows);
        const double v = (j >= c->cols / 2 ? -(double)(c->cols - j) : (double)j) /
                         (c->cols * c->delta_cols);
        const double tran = c->dist * reg * (u * u + v * v) * 4 * M_PI * M_PI + 1;
        f[(size_t)i * c->cols + j] = make_float2((float)(1.0 / (tran * N)), 0.f);
      }
    if (!c->fac_pag && dev_alloc(&c->fac_pag, c->npix)) return -1;
    HIP_OK(hipMemcpy(c->fac_pag, f.data(), f.size() * sizeof(float2), hipMemcpyHostToDevice));
    c->pag_dob = delta_over_beta;
    c->pag_k = wavenumber;
  }
  if (alloc_nms_state(c)) return -1;
  const size_t cb = (size_t)c->nall * sizeof(float2), db = (size_t)c->nall * sizeof(double);
  float2* yb = c->b + c->nall;  // second half of the transform buffer
  HIP_OK(hipMemcpyAsync(yb, y, cb, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->wold, D, cb, hipMemcpyHostToDevice, c->stream));
  const unsigned eb = ew_blocks(c->nall);
  hipLaunchKernelGGL(k_pag_pre, dim3(eb), dim3(kThreads), 0, c->stream, yb, c->wold, c->nall,
                     c->b);
  if (filter_dev(c, c->b, c->nv, c->fac_pag)) return -1;
  hipLaunchKernelGGL(k_pag_post, dim3(eb), dim3(kThreads), 0, c->stream, c->b, c->nall,
                     delta_over_beta, c->zr, c->zi);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(zr, c->zr, db, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(zi, c->zi, db, hipMemcpyDeviceToHost, c->stream));
  return collect_events(c);
}

int pcmbir_hip_set_option(pcmbir_hip_ctx* c, const char* key, double value) {
  if (!c || !key) return fail("null argument");
  if (!strcmp(key, "fft_reg")) {
    c->fft_reg = value != 0.0;
    return 0;
  }
  if (!strcmp(key, "fft_col_threads")) {
    const int v = (int)value;
    if ((double)v != value || v < 64 || v > 1024 || (v & (v - 1)))
      return fail("fft_col_threads must be a power of two in [64, 1024]");
    c->col_threads = v;
    return 0;
  }
  return fail(std::string("unknown option ") + key);
}

int pcmbir_hip_fft_time(pcmbir_hip_ctx* c, double* ms, int64_t* launches) {
  if (!c) return fail("null context");
  if (collect_events(c)) return -1;
  if (ms) *ms = c->fft_ms;
  if (launches) *launches = c->fft_launches;
  c->fft_ms = 0.0;
  c->fft_launches = 0;
  return 0;
}

}  // extern "C"
