// unet_hip.hip -- learned warm start for the MBIR solver: a 3-level 3D U-Net
// (channels c / 2c / 4c, two 3x3x3 convolutions per level, average pooling,
// nearest up-sampling, concatenated skips, residual output) trained and run
// on the GPU with BF16 MFMA kernels and FP32 accumulation.
//
// The network has no counterpart in the reference: it maps the normalised
// back-projection  sum(w a y) / sum(w a^2)  of every voxel to the converged
// MBIR image, and the MBIR solver then starts from its output instead of from
// zero (it always runs to convergence).
//
// Layout: activations and their gradients are BF16, channels last,
// [n][z][y][x][channel stride]; a layer reads / writes a channel range
// (offset, count) of a buffer, so a concatenation is two layers writing the
// two halves of one buffer.  Weights: FP32 master copy [cout][27][cin]
// (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), packed BF16 for the forward GEMM
// [cout16][kpad] and, with taps flipped and channels swapped, for the
// backward-data GEMM.
//
// Kernels
//   k_conv       implicit-GEMM 3x3x3 convolution, v_mfma_f32_16x16x32_bf16:
//                one wavefront = 16-64 voxels x all output channels, each
//                weight fragment feeding every voxel fragment; forward
//                (bias, ReLU) and backward-data (flipped weights)
//   k_conv_dw    weight gradient: (16 out channels) x (16 im2col columns)
//                per wavefront, reduced over a voxel chunk with MFMA, FP32
//                atomics into the gradient
//   k_pool / k_pool_bwd, k_up / k_up_bwd, k_relu_bwd, k_bias_grad, k_mse,
//   k_adam, k_pack   elementwise / reduction helpers
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "unet_hip.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

thread_local std::string g_err;
int fail(const std::string& m) {
  g_err = m;
  return -1;
}
#define HIP_OK(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      return fail(std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__host__ __device__ inline u16 f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

inline int rup(int a, int b) { return (a + b - 1) / b * b; }

// A channel range of a channels-last BF16 buffer.
struct Chan {
  u16* p;
  int cs, co;  // channel stride, first channel
};

struct ConvArgs {
  const u16* in;
  int cs_in, co_in;
  const u16* w;  // [nt*16][kpad]
  const float* bias;
  u16* out;      // BF16 output channel range (may be null)
  int cs_out, co_out;
  float* out32;  // FP32 output [v][cout] (may be null)
  int N, D, H, W;
  int cin_pad, kpad, cout, relu;
};

// Output channel tiles of 16 per wavefront: NT; 16-voxel fragments per
// wavefront: MV (each weight fragment feeds MV MFMAs).  Lane l: voxel row
// l % 16 of an A fragment with 8 consecutive k (one tap, 8 channels: a
// 16-byte load), output channel l % 16 of the B operand; a result holds
// output channel l % 16 of voxels 4 (l / 16) + r of its fragment.  A
// fragment past the last voxel feeds zeros and stores nothing.
template <int NT, int MV>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const long long V = (long long)a.N * a.D * a.H * a.W;
  const long long v0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * (16 * MV);
  if (v0 >= V) return;
  const int m = lane & 15, kg = lane >> 4;
  int x[MV], y[MV], z[MV];
  long long n[MV];
  bool in[MV];
#pragma unroll
  for (int f = 0; f < MV; ++f) {
    const long long v = v0 + f * 16 + m;
    in[f] = v0 + f * 16 < V;
    x[f] = (int)(v % a.W);
    long long r = v / a.W;
    y[f] = (int)(r % a.H);
    r /= a.H;
    z[f] = (int)(r % a.D);
    n[f] = r / a.D;
  }
  floatx4 acc[MV][NT];
#pragma unroll
  for (int f = 0; f < MV; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < a.kpad; k0 += 32) {
    const int k = k0 + kg * 8;
    const int tap = k / a.cin_pad, ci = k - tap * a.cin_pad;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    bf16x8 av[MV], bv[NT];
#pragma unroll
    for (int f = 0; f < MV; ++f) {
      uint4 raw = make_uint4(0u, 0u, 0u, 0u);
      const int zz = z[f] + dz, yy = y[f] + dy, xx = x[f] + dx;
      if (in[f] && tap < 27 && zz >= 0 && zz < a.D && yy >= 0 && yy < a.H && xx >= 0 &&
          xx < a.W) {
        const long long vv = ((n[f] * a.D + zz) * a.H + yy) * a.W + xx;
        raw = *reinterpret_cast<const uint4*>(a.in + vv * a.cs_in + a.co_in + ci);
      }
      av[f] = __builtin_bit_cast(bf16x8, raw);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
      bv[t] = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(a.w + (size_t)(t * 16 + m) * a.kpad + k));
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int f = 0; f < MV; ++f)
        acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[f], bv[t], acc[f][t], 0, 0, 0);
  }
#pragma unroll
  for (int f = 0; f < MV; ++f) {
    if (!in[f]) continue;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = t * 16 + m;
      if (co >= a.cout) continue;
      const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long long vo = v0 + f * 16 + kg * 4 + q;
        float o = acc[f][t][q] + b;
        if (a.relu) o = fmaxf(o, 0.f);
        if (a.out) a.out[vo * a.cs_out + a.co_out + co] = f2bf(o);
        if (a.out32) a.out32[vo * a.cout + co] = o;
      }
    }
  }
}

struct DwArgs {
  const u16* g;  // output gradient channel range (the GEMM's M side)
  int cs_g, co_g;
  const u16* in;  // layer input channel range
  int cs_in, co_in;
  float* dw;  // [cout][27 * cin_pad]  (FP32, accumulated)
  int N, D, H, W;
  int cin_pad, cout, kcols;  // kcols = 27 * cin_pad
  int chunk;                 // voxels per wavefront (multiple of 32)
};

// dW[co][k] = sum_v g[v][co] im2col(in)[v][k]: a wavefront takes out channels
// 16 co-tile .. +15 (A rows) and im2col columns 16 k-tile .. +15 (B columns)
// over `chunk` voxels, 32 per MFMA; blockIdx.x = voxel chunk, y = group of 4
// k tiles (one per wavefront), z = co tile.
__global__ __launch_bounds__(256) void k_conv_dw(DwArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const int ktile = blockIdx.y * 4 + wv;
  const int kcol = ktile * 16 + m;
  if (ktile * 16 >= a.kcols) return;
  const int co = blockIdx.z * 16 + m;
  const long long V = (long long)a.N * a.D * a.H * a.W;
  const long long vb = (long long)blockIdx.x * a.chunk;
  const long long ve = vb + a.chunk < V ? vb + a.chunk : V;
  const bool kok = kcol < a.kcols;
  const int tap = kok ? kcol / a.cin_pad : 0, ci = kok ? kcol - tap * a.cin_pad : 0;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
  for (long long s = vb; s < ve; s += 32) {
    u16 ga[8], ib[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const long long v = s + kg * 8 + j;
      ga[j] = 0;
      ib[j] = 0;
      if (v < ve) {
        if (co < a.cout) ga[j] = a.g[v * a.cs_g + a.co_g + co];
        if (kok) {
          const int x = (int)(v % a.W);
          long long r = v / a.W;
          const int y = (int)(r % a.H);
          r /= a.H;
          const int z = (int)(r % a.D);
          const long long n = r / a.D;
          const int zz = z + dz, yy = y + dy, xx = x + dx;
          if (zz >= 0 && zz < a.D && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
            ib[j] = a.in[(((n * a.D + zz) * a.H + yy) * a.W + xx) * a.cs_in + a.co_in + ci];
        }
      }
    }
    bf16x8 av, bv;
    av = __builtin_bit_cast(bf16x8, make_uint4(ga[0] | (uint32_t)ga[1] << 16,
                                               ga[2] | (uint32_t)ga[3] << 16,
                                               ga[4] | (uint32_t)ga[5] << 16,
                                               ga[6] | (uint32_t)ga[7] << 16));
    bv = __builtin_bit_cast(bf16x8, make_uint4(ib[0] | (uint32_t)ib[1] << 16,
                                               ib[2] | (uint32_t)ib[3] << 16,
                                               ib[4] | (uint32_t)ib[5] << 16,
                                               ib[6] | (uint32_t)ib[7] << 16));
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  // result: k column kcol = ktile*16 + (lane & 15), out channels 16 co-tile + 4 kg + q
  if (!kok) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = blockIdx.z * 16 + kg * 4 + q;
    if (c < a.cout && acc[q] != 0.f) atomicAdd(&a.dw[(size_t)c * a.kcols + kcol], acc[q]);
  }
}

// 2x2x2 average pooling: out (D/2, H/2, W/2) channel range := mean of 8.
__global__ void k_pool(const u16* in, int cs_in, int co_in, u16* out, int cs_out, int co_out,
                       int C, int N, int D, int H, int W) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
  const long long total = (long long)N * D2 * H2 * W2 * C;
  if (i >= total) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int x = (int)(v % W2);
  v /= W2;
  const int y = (int)(v % H2);
  v /= H2;
  const int z = (int)(v % D2);
  const long long n = v / D2;
  float s = 0.f;
  for (int d = 0; d < 8; ++d) {
    const long long src = ((n * D + 2 * z + (d >> 2)) * H + 2 * y + ((d >> 1) & 1)) * W + 2 * x + (d & 1);
    s += bf2f(in[src * cs_in + co_in + c]);
  }
  out[(((n * D2 + z) * H2 + y) * W2 + x) * cs_out + co_out + c] = f2bf(s * 0.125f);
}

// gradient of the pooling: gin (D, H, W) channel range += gout / 8
__global__ void k_pool_bwd(const u16* gout, int cs_o, int co_o, u16* gin, int cs_i, int co_i,
                           int C, int N, int D, int H, int W) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)N * D * H * W * C;
  if (i >= total) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int x = (int)(v % W);
  v /= W;
  const int y = (int)(v % H);
  v /= H;
  const int z = (int)(v % D);
  const long long n = v / D;
  const long long p = ((n * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2;
  const long long dst = (((n * D + z) * H + y) * W + x) * cs_i + co_i + c;
  gin[dst] = f2bf(bf2f(gin[dst]) + 0.125f * bf2f(gout[p * cs_o + co_o + c]));
}

// nearest 2x up-sampling: out (2D, 2H, 2W) channel range := in (D, H, W);
// one thread per 8 channels (16-byte accesses; channel counts, strides and
// offsets are multiples of 8)
__global__ void k_up(const u16* in, int cs_in, int co_in, u16* out, int cs_out, int co_out, int C,
                     int N, int D, int H, int W) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int D2 = 2 * D, H2 = 2 * H, W2 = 2 * W;
  const int C8 = C / 8;
  const long long total = (long long)N * D2 * H2 * W2 * C8;
  if (i >= total) return;
  const int c = (int)(i % C8) * 8;
  long long v = i / C8;
  const int x = (int)(v % W2);
  v /= W2;
  const int y = (int)(v % H2);
  v /= H2;
  const int z = (int)(v % D2);
  const long long n = v / D2;
  const long long src = ((n * D + z / 2) * H + y / 2) * W + x / 2;
  *reinterpret_cast<uint4*>(out + (((n * D2 + z) * H2 + y) * W2 + x) * cs_out + co_out + c) =
      *reinterpret_cast<const uint4*>(in + src * cs_in + co_in + c);
}

// gradient of the up-sampling: gin (D, H, W) := sum of its 8 children
__global__ void k_up_bwd(const u16* gout, int cs_o, int co_o, u16* gin, int cs_i, int co_i, int C,
                         int N, int D, int H, int W) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)N * D * H * W * C;
  if (i >= total) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int x = (int)(v % W);
  v /= W;
  const int y = (int)(v % H);
  v /= H;
  const int z = (int)(v % D);
  const long long n = v / D;
  float s = 0.f;
  for (int d = 0; d < 8; ++d) {
    const long long src =
        ((n * 2 * D + 2 * z + (d >> 2)) * 2 * H + 2 * y + ((d >> 1) & 1)) * 2 * W + 2 * x + (d & 1);
    s += bf2f(gout[src * cs_o + co_o + c]);
  }
  gin[(((n * D + z) * H + y) * W + x) * cs_i + co_i + c] = f2bf(s);
}

// g := g where act > 0, else 0 (ReLU gradient through the stored output)
__global__ void k_relu_bwd(u16* g, int cs_g, int co_g, const u16* act, int cs_a, int co_a, int C,
                           long long V) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  if (!(bf2f(act[v * cs_a + co_a + c]) > 0.f)) g[v * cs_g + co_g + c] = 0;
}

// db[c] += sum over the block's voxels of g[v][c]
__global__ void k_bias_grad(const u16* g, int cs, int co, int C, long long V, float* db) {
  __shared__ float red[256];
  const int c = blockIdx.y;
  float s = 0.f;
  for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < V; v += (long long)gridDim.x * 256)
    s += bf2f(g[v * cs + co + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&db[c], red[0]);
}

// pred = net + x (input channel 0 of cin); grad channel 0 := 2 (pred - t) / V
// (others 0); partial sums of (pred - t)^2 per block
__global__ void k_mse(const float* net, const float* x, int cin, const float* t, long long V,
                      u16* g, int cs_g, double* partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < V; v += (long long)gridDim.x * 256) {
    const float d = net[v] + x[v * cin] - t[v];
    s += (double)d * d;
    for (int c = 0; c < cs_g; ++c) g[v * cs_g + c] = c == 0 ? f2bf(2.f * d / (float)V) : (u16)0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// y[v] += x[v][0] (the residual: input channel 0 of cin)
__global__ void k_add(float* y, const float* x, int cin, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i * cin];
}

// x (FP32, [v][cin]) into channels 0 .. cin-1 of a BF16 buffer of stride cs
// (the others 0)
__global__ void k_in_pack(const float* x, int cin, long long V, u16* out, int cs) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * cs) return;
  const int c = (int)(i % cs);
  out[i] = c < cin ? f2bf(x[(i / cs) * cin + c]) : (u16)0;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 master [cout][27][cin] -> forward [cout16][kpad] (k = tap*cin_pad + ci)
// and backward-data [cin16][kpadT] (k = (26 - tap)*cout_pad + co)
__global__ void k_pack(const float* w, int cin, int cout, int cin_pad, int kpad, int cout_pad,
                       int kpadT, u16* wf, u16* wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cout * 27 * cin) return;
  const int ci = i % cin, tap = (i / cin) % 27, co = i / (27 * cin);
  const u16 h = f2bf(w[i]);
  wf[(size_t)co * kpad + tap * cin_pad + ci] = h;
  wb[(size_t)ci * kpadT + (26 - tap) * cout_pad + co] = h;
}

// One wavefront: D = A B with A 16x32, B 32x16 (row-major FP32 in, rounded
// to BF16), through the lane layout k_conv and k_conv_dw use.
__global__ void k_mfma_probe(const float* A, const float* B, float* Dm) {
  const int lane = threadIdx.x, m = lane & 15, kg = lane >> 4;
  u16 a[8], b[8];
  for (int j = 0; j < 8; ++j) {
    a[j] = f2bf(A[m * 32 + kg * 8 + j]);
    b[j] = f2bf(B[(kg * 8 + j) * 16 + m]);
  }
  const bf16x8 av = __builtin_bit_cast(
      bf16x8, make_uint4(a[0] | (uint32_t)a[1] << 16, a[2] | (uint32_t)a[3] << 16,
                         a[4] | (uint32_t)a[5] << 16, a[6] | (uint32_t)a[7] << 16));
  const bf16x8 bv = __builtin_bit_cast(
      bf16x8, make_uint4(b[0] | (uint32_t)b[1] << 16, b[2] | (uint32_t)b[3] << 16,
                         b[4] | (uint32_t)b[5] << 16, b[6] | (uint32_t)b[7] << 16));
  floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  for (int q = 0; q < 4; ++q) Dm[(kg * 4 + q) * 16 + m] = acc[q];
}

template <typename T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  HIP_OK(hipMalloc((void**)p, (n ? n : 1) * sizeof(T)));
  HIP_OK(hipMemset(*p, 0, (n ? n : 1) * sizeof(T)));
  return 0;
}

unsigned blocks_for(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// One 3x3x3 convolution layer.
struct Layer {
  int cin = 0, cout = 0, relu = 1;
  int cin_pad = 0, kpad = 0, nt = 0;       // forward GEMM
  int cout_pad = 0, kpadT = 0, ntT = 0;    // backward-data GEMM
  float *w = nullptr, *b = nullptr;        // FP32 master
  float *gw = nullptr, *gb = nullptr;      // gradients
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
  u16 *wf = nullptr, *wb = nullptr;
  size_t nw() const { return (size_t)cout * 27 * cin; }
  int init(int ci, int co, int r) {
    cin = ci;
    cout = co;
    relu = r;
    cin_pad = rup(ci, 8);
    kpad = rup(27 * cin_pad, 32);
    nt = rup(co, 16) / 16;
    cout_pad = rup(co, 8);
    kpadT = rup(27 * cout_pad, 32);
    ntT = rup(ci, 16) / 16;
    if (dalloc(&w, nw()) || dalloc(&b, cout) || dalloc(&gw, (size_t)cout * 27 * cin_pad) ||
        dalloc(&gb, cout) || dalloc(&mw, nw()) || dalloc(&vw, nw()) || dalloc(&mb, cout) ||
        dalloc(&vb, cout) || dalloc(&wf, (size_t)nt * 16 * kpad) ||
        dalloc(&wb, (size_t)ntT * 16 * kpadT))
      return -1;
    return 0;
  }
  void release() {
    void* ps[] = {w, b, gw, gb, mw, vw, mb, vb, wf, wb};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  int pack(hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((long long)nw())), dim3(256), 0, s, w, cin, cout,
                       cin_pad, kpad, cout_pad, kpadT, wf, wb);
    HIP_OK(hipGetLastError());
    return 0;
  }
};

int launch_conv(hipStream_t s, const Chan& in, const u16* w, int cin_pad, int kpad, int nt,
                const float* bias, const Chan* out, float* out32, int cout, int relu, int N, int D,
                int H, int W) {
  ConvArgs a;
  a.in = in.p;
  a.cs_in = in.cs;
  a.co_in = in.co;
  a.w = w;
  a.bias = bias;
  a.out = out ? out->p : nullptr;
  a.cs_out = out ? out->cs : 0;
  a.co_out = out ? out->co : 0;
  a.out32 = out32;
  a.N = N;
  a.D = D;
  a.H = H;
  a.W = W;
  a.cin_pad = cin_pad;
  a.kpad = kpad;
  a.cout = cout;
  a.relu = relu;
  const long long V = (long long)N * D * H * W;
  if (V % 16) return fail("voxel count must be a multiple of 16");
  if (in.cs % 8 || in.co % 8 || (out && (out->cs < out->co + cout)))
    return fail("channel strides and offsets must be multiples of 8");
  // fragments per wavefront: 4 for up to 32 output channels, 2 for 48-64,
  // 1 above (accumulators MV x NT x 4 VGPRs)
  const int mv = nt <= 2 ? 4 : (nt <= 4 ? 2 : 1);
  const dim3 g(blocks_for((V + 16 * mv - 1) / (16 * mv), 4));
  switch (nt) {
    case 1: hipLaunchKernelGGL((k_conv<1, 4>), g, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_conv<2, 4>), g, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_conv<3, 2>), g, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_conv<4, 2>), g, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_conv<6, 1>), g, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_conv<8, 1>), g, dim3(256), 0, s, a); break;
    default: return fail("unsupported channel count");
  }
  HIP_OK(hipGetLastError());
  return 0;
}

int launch_dw(hipStream_t s, const Chan& g, const Chan& in, float* dw, int cin_pad, int cout,
              int N, int D, int H, int W) {
  DwArgs a;
  a.g = g.p;
  a.cs_g = g.cs;
  a.co_g = g.co;
  a.in = in.p;
  a.cs_in = in.cs;
  a.co_in = in.co;
  a.dw = dw;
  a.N = N;
  a.D = D;
  a.H = H;
  a.W = W;
  a.cin_pad = cin_pad;
  a.cout = cout;
  a.kcols = 27 * cin_pad;
  const long long V = (long long)N * D * H * W;
  a.chunk = 2048;
  const int ktiles = (a.kcols + 15) / 16;
  const dim3 grid((unsigned)((V + a.chunk - 1) / a.chunk), (unsigned)((ktiles + 3) / 4),
                  (unsigned)((cout + 15) / 16));
  hipLaunchKernelGGL(k_conv_dw, grid, dim3(256), 0, s, a);
  HIP_OK(hipGetLastError());
  return 0;
}

}  // namespace

// ------------------------------------------------------------------ network
struct unet_ctx {
  hipStream_t s = nullptr;
  int c = 16, cin = 1;  // base channels, input channels
  Layer L[11];
  // shape the activation buffers hold
  int N = 0, D = 0, H = 0, W = 0;
  long long V1 = 0, V2 = 0, V3 = 0;
  u16 *X0 = nullptr, *A1 = nullptr, *C1 = nullptr, *P1 = nullptr, *A2 = nullptr, *C2 = nullptr,
      *P2 = nullptr, *A3 = nullptr, *B3 = nullptr, *A4 = nullptr, *B4 = nullptr, *A5 = nullptr,
      *B5 = nullptr;
  u16 *GO = nullptr, *GB5 = nullptr, *GA5 = nullptr, *GC1 = nullptr, *GB4 = nullptr,
      *GA4 = nullptr, *GC2 = nullptr, *GB3 = nullptr, *GA3 = nullptr, *GP2 = nullptr,
      *GA2 = nullptr, *GP1 = nullptr, *GA1 = nullptr;
  float *x32 = nullptr, *t32 = nullptr, *o32 = nullptr;
  double* partial = nullptr;
  int step = 0;
  std::vector<void*> acts;
  void free_acts() {
    for (void* p : acts)
      if (p) (void)hipFree(p);
    acts.clear();
  }
};

namespace {

constexpr int kLossBlocks = 256;

int alloc_act(unet_ctx* c, u16** p, long long n) {
  if (dalloc(p, (size_t)n)) return -1;
  c->acts.push_back(*p);
  return 0;
}

int set_shape(unet_ctx* c, int N, int D, int H, int W) {
  if (N == c->N && D == c->D && H == c->H && W == c->W) return 0;
  if (N < 1 || D < 4 || H < 4 || W < 64 || D % 4 || H % 4 || W % 64)
    return fail("volume must have D, H divisible by 4 and W divisible by 64");
  HIP_OK(hipStreamSynchronize(c->s));
  c->free_acts();
  c->N = N;
  c->D = D;
  c->H = H;
  c->W = W;
  const long long V1 = (long long)N * D * H * W, V2 = V1 / 8, V3 = V1 / 64;
  c->V1 = V1;
  c->V2 = V2;
  c->V3 = V3;
  const int c1 = c->c, c2 = 2 * c->c, c3 = 4 * c->c;
  u16** bufs1[] = {&c->A1, &c->A5, &c->B5, &c->GB5, &c->GA5, &c->GA1};
  for (u16** b : bufs1)
    if (alloc_act(c, b, V1 * c1)) return -1;
  if (alloc_act(c, &c->X0, V1 * 8) || alloc_act(c, &c->GO, V1 * 8) ||
      alloc_act(c, &c->C1, V1 * (c2 + c1)) || alloc_act(c, &c->GC1, V1 * (c2 + c1)))
    return -1;
  u16** bufs2[] = {&c->A2, &c->A4, &c->B4, &c->GB4, &c->GA4, &c->GA2};
  for (u16** b : bufs2)
    if (alloc_act(c, b, V2 * c2)) return -1;
  if (alloc_act(c, &c->P1, V2 * c1) || alloc_act(c, &c->GP1, V2 * c1) ||
      alloc_act(c, &c->C2, V2 * (c3 + c2)) || alloc_act(c, &c->GC2, V2 * (c3 + c2)))
    return -1;
  u16** bufs3[] = {&c->A3, &c->B3, &c->GB3, &c->GA3};
  for (u16** b : bufs3)
    if (alloc_act(c, b, V3 * c3)) return -1;
  if (alloc_act(c, &c->P2, V3 * c2) || alloc_act(c, &c->GP2, V3 * c2)) return -1;
  float** f32s[] = {&c->x32, &c->t32, &c->o32};
  for (float** b : f32s) {
    if (dalloc(b, (size_t)V1 * (b == &c->x32 ? c->cin : 1))) return -1;
    c->acts.push_back(*b);
  }
  return 0;
}

int conv_layer(unet_ctx* c, int li, const Chan& in, const Chan* out, float* out32, int lev) {
  const Layer& l = c->L[li];
  const int sh = lev;  // spatial halving per level
  return launch_conv(c->s, in, l.wf, l.cin_pad, l.kpad, l.nt, l.b, out, out32, l.cout, l.relu,
                     c->N, c->D >> sh, c->H >> sh, c->W >> sh);
}

int forward(unet_ctx* c) {
  const int c1 = c->c, c2 = 2 * c->c, c3 = 4 * c->c;
  const int N = c->N, D = c->D, H = c->H, W = c->W;
  const Chan X0{c->X0, 8, 0}, A1{c->A1, c1, 0}, S1{c->C1, c2 + c1, c2}, U1{c->C1, c2 + c1, 0},
      C1{c->C1, c2 + c1, 0}, P1{c->P1, c1, 0}, A2{c->A2, c2, 0}, S2{c->C2, c3 + c2, c3},
      U2{c->C2, c3 + c2, 0}, C2{c->C2, c3 + c2, 0}, P2{c->P2, c2, 0}, A3{c->A3, c3, 0},
      B3{c->B3, c3, 0}, A4{c->A4, c2, 0}, B4{c->B4, c2, 0}, A5{c->A5, c1, 0}, B5{c->B5, c1, 0};
  hipLaunchKernelGGL(k_in_pack, dim3(blocks_for(c->V1 * 8)), dim3(256), 0, c->s, c->x32, c->cin,
                     c->V1, c->X0, 8);
  if (conv_layer(c, 0, X0, &A1, nullptr, 0) || conv_layer(c, 1, A1, &S1, nullptr, 0)) return -1;
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(c->V2 * c1)), dim3(256), 0, c->s, S1.p, S1.cs, S1.co,
                     P1.p, P1.cs, P1.co, c1, N, D, H, W);
  if (conv_layer(c, 2, P1, &A2, nullptr, 1) || conv_layer(c, 3, A2, &S2, nullptr, 1)) return -1;
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(c->V3 * c2)), dim3(256), 0, c->s, S2.p, S2.cs, S2.co,
                     P2.p, P2.cs, P2.co, c2, N, D / 2, H / 2, W / 2);
  if (conv_layer(c, 4, P2, &A3, nullptr, 2) || conv_layer(c, 5, A3, &B3, nullptr, 2)) return -1;
  hipLaunchKernelGGL(k_up, dim3(blocks_for(c->V2 * c3 / 8)), dim3(256), 0, c->s, B3.p, B3.cs, B3.co,
                     U2.p, U2.cs, U2.co, c3, N, D / 4, H / 4, W / 4);
  if (conv_layer(c, 6, C2, &A4, nullptr, 1) || conv_layer(c, 7, A4, &B4, nullptr, 1)) return -1;
  hipLaunchKernelGGL(k_up, dim3(blocks_for(c->V1 * c2 / 8)), dim3(256), 0, c->s, B4.p, B4.cs, B4.co,
                     U1.p, U1.cs, U1.co, c2, N, D / 2, H / 2, W / 2);
  if (conv_layer(c, 8, C1, &A5, nullptr, 0) || conv_layer(c, 9, A5, &B5, nullptr, 0) ||
      conv_layer(c, 10, B5, nullptr, c->o32, 0))
    return -1;
  HIP_OK(hipGetLastError());
  return 0;
}

// gradients of one layer: relu mask on g (when the layer has one), dW, db,
// and (when gin) the input gradient
int layer_bwd(unet_ctx* c, int li, const Chan& g, const Chan& act_out, const Chan& in,
              const Chan* gin, int lev) {
  const Layer& l = c->L[li];
  const int N = c->N, D = c->D >> lev, H = c->H >> lev, W = c->W >> lev;
  const long long V = (long long)N * D * H * W;
  if (l.relu)
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(V * l.cout)), dim3(256), 0, c->s, g.p, g.cs,
                       g.co, act_out.p, act_out.cs, act_out.co, l.cout, V);
  if (launch_dw(c->s, g, in, l.gw, l.cin_pad, l.cout, N, D, H, W)) return -1;
  hipLaunchKernelGGL(k_bias_grad, dim3(64, l.cout), dim3(256), 0, c->s, g.p, g.cs, g.co, l.cout, V,
                     l.gb);
  if (gin && launch_conv(c->s, g, l.wb, l.cout_pad, l.kpadT, l.ntT, nullptr, gin, nullptr, l.cin,
                         0, N, D, H, W))
    return -1;
  HIP_OK(hipGetLastError());
  return 0;
}

int backward(unet_ctx* c, double* loss) {
  const int c1 = c->c, c2 = 2 * c->c, c3 = 4 * c->c;
  const int N = c->N, D = c->D, H = c->H, W = c->W;
  const Chan X0{c->X0, 8, 0}, A1{c->A1, c1, 0}, S1{c->C1, c2 + c1, c2}, C1{c->C1, c2 + c1, 0},
      P1{c->P1, c1, 0}, A2{c->A2, c2, 0}, S2{c->C2, c3 + c2, c3}, C2{c->C2, c3 + c2, 0},
      P2{c->P2, c2, 0}, A3{c->A3, c3, 0}, B3{c->B3, c3, 0}, A4{c->A4, c2, 0}, B4{c->B4, c2, 0},
      A5{c->A5, c1, 0}, B5{c->B5, c1, 0};
  const Chan GO{c->GO, 8, 0}, GB5{c->GB5, c1, 0}, GA5{c->GA5, c1, 0}, GC1{c->GC1, c2 + c1, 0},
      GU1{c->GC1, c2 + c1, 0}, GS1{c->GC1, c2 + c1, c2}, GB4{c->GB4, c2, 0}, GA4{c->GA4, c2, 0},
      GC2{c->GC2, c3 + c2, 0}, GU2{c->GC2, c3 + c2, 0}, GS2{c->GC2, c3 + c2, c3},
      GB3{c->GB3, c3, 0}, GA3{c->GA3, c3, 0}, GP2{c->GP2, c2, 0}, GA2{c->GA2, c2, 0},
      GP1{c->GP1, c1, 0}, GA1{c->GA1, c1, 0};
  for (auto& l : c->L) {
    HIP_OK(hipMemsetAsync(l.gw, 0, (size_t)l.cout * 27 * l.cin_pad * sizeof(float), c->s));
    HIP_OK(hipMemsetAsync(l.gb, 0, (size_t)l.cout * sizeof(float), c->s));
  }
  hipLaunchKernelGGL(k_mse, dim3(kLossBlocks), dim3(256), 0, c->s, c->o32, c->x32, c->cin,
                     c->t32, c->V1, c->GO, 8, c->partial);
  if (layer_bwd(c, 10, GO, GO, B5, &GB5, 0) || layer_bwd(c, 9, GB5, B5, A5, &GA5, 0) ||
      layer_bwd(c, 8, GA5, A5, C1, &GC1, 0))
    return -1;
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(c->V2 * c2)), dim3(256), 0, c->s, GU1.p, GU1.cs,
                     GU1.co, GB4.p, GB4.cs, GB4.co, c2, N, D / 2, H / 2, W / 2);
  if (layer_bwd(c, 7, GB4, B4, A4, &GA4, 1) || layer_bwd(c, 6, GA4, A4, C2, &GC2, 1)) return -1;
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(c->V3 * c3)), dim3(256), 0, c->s, GU2.p, GU2.cs,
                     GU2.co, GB3.p, GB3.cs, GB3.co, c3, N, D / 4, H / 4, W / 4);
  if (layer_bwd(c, 5, GB3, B3, A3, &GA3, 2) || layer_bwd(c, 4, GA3, A3, P2, &GP2, 2)) return -1;
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(c->V2 * c2)), dim3(256), 0, c->s, GP2.p, GP2.cs,
                     GP2.co, GS2.p, GS2.cs, GS2.co, c2, N, D / 2, H / 2, W / 2);
  if (layer_bwd(c, 3, GS2, S2, A2, &GA2, 1) || layer_bwd(c, 2, GA2, A2, P1, &GP1, 1)) return -1;
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(c->V1 * c1)), dim3(256), 0, c->s, GP1.p, GP1.cs,
                     GP1.co, GS1.p, GS1.cs, GS1.co, c1, N, D, H, W);
  if (layer_bwd(c, 1, GS1, S1, A1, &GA1, 0) || layer_bwd(c, 0, GA1, A1, X0, nullptr, 0)) return -1;
  (void)GC1;
  (void)GC2;
  (void)C1;
  (void)C2;
  if (loss) {
    std::vector<double> part(kLossBlocks);
    HIP_OK(hipMemcpyAsync(part.data(), c->partial, kLossBlocks * sizeof(double),
                          hipMemcpyDeviceToHost, c->s));
    HIP_OK(hipStreamSynchronize(c->s));
    double s = 0.0;
    for (double v : part) s += v;
    *loss = s / (double)c->V1;
  }
  HIP_OK(hipGetLastError());
  return 0;
}

int upload_pair(unet_ctx* c, const float* x, const float* t) {
  HIP_OK(hipMemcpyAsync(c->x32, x, c->V1 * c->cin * sizeof(float), hipMemcpyHostToDevice, c->s));
  if (t) HIP_OK(hipMemcpyAsync(c->t32, t, c->V1 * sizeof(float), hipMemcpyHostToDevice, c->s));
  return 0;
}

}  // namespace

extern "C" {

const char* unet_last_error(void) { return g_err.c_str(); }

int unet_create(unet_ctx** out, int32_t channels, int32_t in_channels, uint32_t seed) {
  if (!out) return fail("null argument");
  *out = nullptr;
  if (channels < 8 || channels > 32 || channels % 8) return fail("channels must be 8, 16, 24 or 32");
  if (in_channels < 1 || in_channels > 8) return fail("in_channels must be in [1, 8]");
  unet_ctx* c = new unet_ctx();
  c->c = channels;
  c->cin = in_channels;
  const int c1 = channels, c2 = 2 * channels, c3 = 4 * channels;
  const int io[11][3] = {{in_channels, c1, 1},  {c1, c1, 1},      {c1, c2, 1}, {c2, c2, 1},
                         {c2, c3, 1}, {c3, c3, 1},      {c3 + c2, c2, 1}, {c2, c2, 1},
                         {c2 + c1, c1, 1}, {c1, c1, 1}, {c1, 1, 0}};
  if (hipStreamCreate(&c->s) != hipSuccess) {
    delete c;
    return fail("hipStreamCreate failed");
  }
  std::mt19937 rng(seed);
  for (int i = 0; i < 11; ++i) {
    if (c->L[i].init(io[i][0], io[i][1], io[i][2])) {
      unet_destroy(c);
      return -1;
    }
    // He-normal initialisation (the output layer scaled down: the network
    // starts near the identity of its residual form)
    const float sd = std::sqrt(2.f / (27.f * io[i][0])) * (i == 10 ? 0.1f : 1.f);
    std::normal_distribution<float> nd(0.f, sd);
    std::vector<float> w(c->L[i].nw());
    for (float& v : w) v = nd(rng);
    if (hipMemcpy(c->L[i].w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess ||
        c->L[i].pack(c->s)) {
      unet_destroy(c);
      return fail("weight upload failed");
    }
  }
  if (dalloc(&c->partial, kLossBlocks)) {
    unet_destroy(c);
    return -1;
  }
  *out = c;
  return 0;
}

void unet_destroy(unet_ctx* c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  c->free_acts();
  for (auto& l : c->L) l.release();
  if (c->partial) (void)hipFree(c->partial);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

int64_t unet_param_count(const unet_ctx* c) {
  if (!c) return -1;
  int64_t n = 0;
  for (const auto& l : c->L) n += (int64_t)l.nw() + l.cout;
  return n;
}

int unet_get_params(unet_ctx* c, float* buf) {
  if (!c || !buf) return fail("null argument");
  for (auto& l : c->L) {
    HIP_OK(hipMemcpy(buf, l.w, l.nw() * sizeof(float), hipMemcpyDeviceToHost));
    buf += l.nw();
    HIP_OK(hipMemcpy(buf, l.b, l.cout * sizeof(float), hipMemcpyDeviceToHost));
    buf += l.cout;
  }
  return 0;
}

int unet_set_params(unet_ctx* c, const float* buf) {
  if (!c || !buf) return fail("null argument");
  HIP_OK(hipStreamSynchronize(c->s));
  for (auto& l : c->L) {
    HIP_OK(hipMemcpy(l.w, buf, l.nw() * sizeof(float), hipMemcpyHostToDevice));
    buf += l.nw();
    HIP_OK(hipMemcpy(l.b, buf, l.cout * sizeof(float), hipMemcpyHostToDevice));
    buf += l.cout;
    if (l.pack(c->s)) return -1;
  }
  HIP_OK(hipStreamSynchronize(c->s));
  return 0;
}

int unet_forward(unet_ctx* c, int32_t N, int32_t D, int32_t H, int32_t W, const float* x,
                 float* y) {
  if (!c || !x || !y) return fail("null argument");
  if (set_shape(c, N, D, H, W) || upload_pair(c, x, nullptr) || forward(c)) return -1;
  hipLaunchKernelGGL(k_add, dim3(blocks_for(c->V1)), dim3(256), 0, c->s, c->o32, c->x32, c->cin,
                     c->V1);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(y, c->o32, c->V1 * sizeof(float), hipMemcpyDeviceToHost, c->s));
  HIP_OK(hipStreamSynchronize(c->s));
  return 0;
}

int unet_gradients(unet_ctx* c, int32_t N, int32_t D, int32_t H, int32_t W, const float* x,
                   const float* t, float* grads, double* loss) {
  if (!c || !x || !t || !grads) return fail("null argument");
  if (set_shape(c, N, D, H, W) || upload_pair(c, x, t) || forward(c) || backward(c, loss))
    return -1;
  std::vector<float> tmp;
  for (auto& l : c->L) {
    tmp.resize((size_t)l.cout * 27 * l.cin_pad);
    HIP_OK(hipMemcpy(tmp.data(), l.gw, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int co = 0; co < l.cout; ++co)
      for (int tap = 0; tap < 27; ++tap)
        for (int ci = 0; ci < l.cin; ++ci)
          *grads++ = tmp[(size_t)co * 27 * l.cin_pad + tap * l.cin_pad + ci];
    HIP_OK(hipMemcpy(grads, l.gb, l.cout * sizeof(float), hipMemcpyDeviceToHost));
    grads += l.cout;
  }
  return 0;
}

int unet_train_step(unet_ctx* c, int32_t N, int32_t D, int32_t H, int32_t W, const float* x,
                    const float* t, float lr, double* loss) {
  if (!c || !x || !t) return fail("null argument");
  if (set_shape(c, N, D, H, W) || upload_pair(c, x, t) || forward(c) || backward(c, loss))
    return -1;
  ++c->step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)c->step), c2 = 1.f - std::pow(b2, (float)c->step);
  for (auto& l : c->L) {
    // gradient of the master layout [cout][27][cin] from [cout][27][cin_pad]
    if (l.cin_pad != l.cin) {
      std::vector<float> g((size_t)l.cout * 27 * l.cin_pad), gc(l.nw());
      HIP_OK(hipMemcpyAsync(g.data(), l.gw, g.size() * sizeof(float), hipMemcpyDeviceToHost, c->s));
      HIP_OK(hipStreamSynchronize(c->s));
      for (size_t i = 0; i < gc.size(); ++i) {
        const size_t ci = i % l.cin, rest = i / l.cin;
        gc[i] = g[rest * l.cin_pad + ci];
      }
      HIP_OK(hipMemcpyAsync(l.gw, gc.data(), gc.size() * sizeof(float), hipMemcpyHostToDevice, c->s));
      HIP_OK(hipStreamSynchronize(c->s));
    }
    hipLaunchKernelGGL(k_adam, dim3(blocks_for((long long)l.nw())), dim3(256), 0, c->s, l.w, l.gw,
                       l.mw, l.vw, (long long)l.nw(), lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(l.cout)), dim3(256), 0, c->s, l.b, l.gb, l.mb, l.vb,
                       (long long)l.cout, lr, b1, b2, eps, c1, c2);
    if (l.pack(c->s)) return -1;
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->s));
  return 0;
}

// ---- single-layer entry points (parity tests of the kernels)
int unet_conv3d(int32_t mode, int32_t N, int32_t D, int32_t H, int32_t W, int32_t cin,
                int32_t cout, const float* in, const float* w, const float* bias, int32_t relu,
                const float* gout, float* out) {
  // mode 0: out [v][cout] = conv(in [v][cin]) (+ bias, ReLU);
  // mode 1: out [v][cin] = input gradient of conv for gout [v][cout];
  // mode 2: out [cout][27][cin] = weight gradient for gout and in.
  // in, gout and w are rounded to BF16 as the network stores them.
  if (mode < 0 || mode > 2 || !out || !w) return fail("bad argument");
  if (cin < 1 || cout < 1 || cin > 128 || cout > 128) return fail("bad channel count");
  const long long V = (long long)N * D * H * W;
  if (V <= 0 || V % 16) return fail("voxel count must be a positive multiple of 16");
  Layer l;
  if (l.init(cin, cout, relu)) return -1;
  hipStream_t s = nullptr;
  HIP_OK(hipStreamCreate(&s));
  int rc = 0;
  u16 *din = nullptr, *dg = nullptr;
  float* o32 = nullptr;
  auto pack_act = [&](const float* src, int C, int cs, u16** dst) -> int {
    std::vector<u16> h((size_t)V * cs, 0);
    for (long long v = 0; v < V; ++v)
      for (int ch = 0; ch < C; ++ch) h[(size_t)v * cs + ch] = f2bf(src[(size_t)v * C + ch]);
    if (dalloc(dst, h.size())) return -1;
    HIP_OK(hipMemcpy(*dst, h.data(), h.size() * sizeof(u16), hipMemcpyHostToDevice));
    return 0;
  };
  do {
    if ((rc = (int)(hipMemcpy(l.w, w, l.nw() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))) {
      rc = fail("upload failed");
      break;
    }
    if (bias && hipMemcpy(l.b, bias, cout * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
      rc = fail("upload failed");
      break;
    }
    if ((rc = l.pack(s))) break;
    if (mode == 0) {
      if (!in) { rc = fail("null input"); break; }
      if ((rc = pack_act(in, cin, l.cin_pad, &din)) || (rc = dalloc(&o32, (size_t)V * cout))) break;
      if ((rc = launch_conv(s, Chan{din, l.cin_pad, 0}, l.wf, l.cin_pad, l.kpad, l.nt,
                            bias ? l.b : nullptr, nullptr, o32, cout, relu, N, D, H, W)))
        break;
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipMemcpy(out, o32, (size_t)V * cout * sizeof(float), hipMemcpyDeviceToHost));
    } else if (mode == 1) {
      if (!gout) { rc = fail("null gradient"); break; }
      if ((rc = pack_act(gout, cout, l.cout_pad, &dg)) || (rc = dalloc(&o32, (size_t)V * cin)))
        break;
      if ((rc = launch_conv(s, Chan{dg, l.cout_pad, 0}, l.wb, l.cout_pad, l.kpadT, l.ntT, nullptr,
                            nullptr, o32, cin, 0, N, D, H, W)))
        break;
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipMemcpy(out, o32, (size_t)V * cin * sizeof(float), hipMemcpyDeviceToHost));
    } else {
      if (!gout || !in) { rc = fail("null input"); break; }
      if ((rc = pack_act(in, cin, l.cin_pad, &din)) || (rc = pack_act(gout, cout, l.cout_pad, &dg)))
        break;
      if ((rc = launch_dw(s, Chan{dg, l.cout_pad, 0}, Chan{din, l.cin_pad, 0}, l.gw, l.cin_pad,
                          cout, N, D, H, W)))
        break;
      HIP_OK(hipStreamSynchronize(s));
      std::vector<float> g((size_t)cout * 27 * l.cin_pad);
      HIP_OK(hipMemcpy(g.data(), l.gw, g.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (int co = 0; co < cout; ++co)
        for (int tap = 0; tap < 27; ++tap)
          for (int ci = 0; ci < cin; ++ci)
            out[((size_t)co * 27 + tap) * cin + ci] = g[((size_t)co * 27 + tap) * l.cin_pad + ci];
    }
  } while (0);
  (void)hipStreamSynchronize(s);
  void* ps[] = {din, dg, o32};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  l.release();
  (void)hipStreamDestroy(s);
  return rc;
}

int unet_mfma_probe(const float* A, const float* B, float* Dm) {
  if (!A || !B || !Dm) return fail("null argument");
  float *dA, *dB, *dD;
  if (dalloc(&dA, 16 * 32) || dalloc(&dB, 32 * 16) || dalloc(&dD, 16 * 16)) return -1;
  HIP_OK(hipMemcpy(dA, A, 16 * 32 * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dB, B, 32 * 16 * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, dA, dB, dD);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(Dm, dD, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dD);
  return 0;
}

}  // extern "C"
