// Learned warm start: a four-level convolutional encoder-decoder (U-Net form, channel widths
// c, 2c, 4c, 8c with c = 32 by default) mapping a log-scaled diffraction amplitude (1 x S x S)
// to the object patch's amplitude and phase (2 x S x S).  No reference counterpart.
//
// Every convolution is an implicit GEMM on the matrix cores: v_mfma_f32_16x16x32_bf16 with BF16
// operands and FP32 accumulation, operands gathered straight from the activation tensors (no
// im2col buffer).  One templated kernel covers the forward pass (M = Cout, N = pixels,
// K = Cin * ks^2), the data gradient (M = Cin, N = pixels, K = Cout * ks^2) and the weight
// gradient (M = Cout, N = Cin * ks^2, K = pixels, split over K into partial tiles summed in a
// fixed order).  Activations are BF16, gradients FP32 (rounded to BF16 where they enter a GEMM),
// master weights FP32 with Adam.  Tensors are NCHW views with a channel stride, so the decoder's
// skip concatenations are channel ranges of one buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../include/ptyk.h"
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

typedef __bf16 bf16;
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

constexpr int GT = 256;  // threads per GEMM block (4 waves)
constexpr int TM = 64, TN = 64, TK = 32;
constexpr int LDP = TK + 8;  // LDS row pitch (bf16) of the A and transposed-B tiles
constexpr int BIAS_PARTS = 256;  // slices of the bias-gradient reduction per channel

// NCHW view: element (b, c, y, x) at p[((b * ctot) + c) * hw + y * w + x]; p already offset to
// the view's first channel.
struct View {
  void* p;
  int c, ctot, h, w;
};

enum { GEMM_FWD = 0, GEMM_BWD_DATA = 1, GEMM_BWD_W = 2 };

struct GemmArgs {
  View x;        // FWD: input activation (bf16); BWD_DATA: dZ (f32); BWD_W: input activation (bf16)
  View y;        // FWD: output (bf16, or f32 when out_f32); BWD_DATA: dX (f32); BWD_W: dZ (f32)
  const bf16* wt;      // weights Cout x Cin x ks x ks (bf16)
  const float* bias;   // FWD
  float* partial;      // BWD_W: chunks x M x N
  int batch, cin, cout, relu, out_f32;
  int M, N, K, kchunk;
  int w_sh, hw_sh;     // log2 of the image width and of the pixels per image (powers of two)
  int vec_a;           // FWD: weight rows 16-byte aligned and K % 8 == 0 (8 bf16 per load)
};

__device__ __forceinline__ bf16 to_bf(float v) { return (bf16)v; }

// A(m, k) for BWD_DATA and BWD_W (FWD reads the weight rows directly)
template <int MODE, int KS>
__device__ __forceinline__ bf16 load_a(const GemmArgs& a, int m, int k) {
  if (m >= a.M || k >= a.K) return to_bf(0.f);
  constexpr int KK2 = KS * KS;
  if (MODE == GEMM_FWD) return a.wt[(long)m * a.K + k];
  if (MODE == GEMM_BWD_DATA) {  // A(ci, (co, d)) = W[co][ci][d]
    const int co = k / KK2, d = k - co * KK2;
    return a.wt[((long)co * a.cin + m) * KK2 + d];
  }
  // BWD_W: A(co, (b, p)) = dZ[b][co][p]
  const int b = k >> a.hw_sh, p = k & ((1 << a.hw_sh) - 1);
  return to_bf(((const float*)a.y.p)[(((long)b * a.y.ctot + m) << a.hw_sh) + p]);
}

// Implicit-GEMM convolution tile: D[m0 .. m0+64) x [n0 .. n0+64) over K in steps of 32.
template <int MODE, int KS>
__global__ __launch_bounds__(GT) void k_conv_gemm(GemmArgs a) {
  constexpr int KK2 = KS * KS, R = KS / 2;
  __shared__ bf16 As[TM * LDP];
  __shared__ bf16 Bt[TN * LDP];  // transposed: Bt[n][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kbeg = MODE == GEMM_BWD_W ? blockIdx.z * a.kchunk : 0;
  const int kend = MODE == GEMM_BWD_W ? min(a.K, kbeg + a.kchunk) : a.K;
  const int wmask = (1 << a.w_sh) - 1, hwmask = (1 << a.hw_sh) - 1;
  // FWD / BWD_DATA: each thread always loads the same B column n0 + tid % 64
  const int bn = n0 + (tid & (TN - 1));
  const bool bn_ok = bn < a.N;
  const int bb = bn >> a.hw_sh, bp = bn & hwmask, by = bp >> a.w_sh, bx = bp & wmask;
  v4f acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    // A tile 64 x 32
    if (MODE == GEMM_FWD && a.vec_a) {
      const int mm = tid >> 2, kk = (tid & 3) * 8;
      const int m = m0 + mm, k = k0 + kk;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (m < a.M && k < a.K) v = *reinterpret_cast<const uint4*>(a.wt + (long)m * a.K + k);
      *reinterpret_cast<uint4*>(&As[mm * LDP + kk]) = v;
    } else {
#pragma unroll
      for (int i = 0; i < TM * TK / GT; ++i) {
        const int e = tid + i * GT;
        const int mm = e / TK, kk = e % TK;
        const int k = k0 + kk;
        As[mm * LDP + kk] = k < kend ? load_a<MODE, KS>(a, m0 + mm, k) : to_bf(0.f);
      }
    }
    // B tile 32 x 64
    if (MODE == GEMM_FWD || MODE == GEMM_BWD_DATA) {
      const View& v = a.x;
#pragma unroll
      for (int i = 0; i < TN * TK / GT; ++i) {
        const int kk = (tid >> 6) + i * (GT / TN);
        const int k = k0 + kk;
        bf16 val = to_bf(0.f);
        if (bn_ok && k < kend) {
          const int ch = k / KK2, d = k - ch * KK2;
          int dy = d / KS - R, dx = d % KS - R;
          if (MODE == GEMM_BWD_DATA) {
            dy = -dy;
            dx = -dx;
          }
          const int sy = by + dy, sx = bx + dx;
          if (sy >= 0 && sy < v.h && sx >= 0 && sx < v.w) {
            const long idx = (((long)bb * v.ctot + ch) << a.hw_sh) + (sy << a.w_sh) + sx;
            val = MODE == GEMM_FWD ? ((const bf16*)v.p)[idx] : to_bf(((const float*)v.p)[idx]);
          }
        }
        Bt[(tid & (TN - 1)) * LDP + kk] = val;
      }
    } else {  // BWD_W: B((b, p), (ci, d)) = x[b][ci][p + d]; consecutive threads along k
      const View& v = a.x;
#pragma unroll
      for (int i = 0; i < TN * TK / GT; ++i) {
        const int e = tid + i * GT;
        const int kk = e % TK, nn = e / TK;
        const int k = k0 + kk, n = n0 + nn;
        bf16 val = to_bf(0.f);
        if (k < kend && n < a.N) {
          const int b = k >> a.hw_sh, p = k & hwmask, yy = p >> a.w_sh, xx = p & wmask;
          const int ci = n / KK2, d = n - ci * KK2;
          const int sy = yy + d / KS - R, sx = xx + d % KS - R;
          if (sy >= 0 && sy < v.h && sx >= 0 && sx < v.w)
            val = ((const bf16*)v.p)[(((long)b * v.ctot + ci) << a.hw_sh) + (sy << a.w_sh) + sx];
        }
        Bt[nn * LDP + kk] = val;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const v8bf va = *reinterpret_cast<const v8bf*>(&As[(wm + i * 16 + (lane & 15)) * LDP + 8 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const v8bf vb = *reinterpret_cast<const v8bf*>(&Bt[(wn + j * 16 + (lane & 15)) * LDP + 8 * (lane >> 4)]);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // epilogue: lane holds D[4 (lane / 16) + r][lane % 16] of each 16 x 16 sub-tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= a.M || n >= a.N) continue;
        float v = acc[i][j][r];
        if (MODE == GEMM_BWD_W) {
          a.partial[((long)blockIdx.z * a.M + m) * a.N + n] = v;
        } else {
          const View& o = a.y;
          const int b = n >> a.hw_sh, p = n & hwmask;
          const long idx = (((long)b * o.ctot + m) << a.hw_sh) + p;
          if (MODE == GEMM_FWD) {
            v += a.bias[m];
            if (a.relu) v = fmaxf(v, 0.f);
            if (a.out_f32) ((float*)o.p)[idx] = v;
            else ((bf16*)o.p)[idx] = to_bf(v);
          } else {
            ((float*)o.p)[idx] = v;
          }
        }
      }
}

// ---- vector-operand form (image width >= 8): each thread moves 8 consecutive pixels of one
// image row per K step -- one aligned 16-byte (or two for FP32) load plus the neighbour element
// a +-1 tap shift needs -- instead of 8 gathered elements.
typedef unsigned short u16;

__device__ __forceinline__ u16 bf_bits(float v) { return __builtin_bit_cast(u16, to_bf(v)); }

// row[x0 + dx .. x0 + dx + 7] (bf16 bits), zero outside [0, w); x0 % 8 == 0, |dx| <= 1
__device__ __forceinline__ void row8_bf16(const u16* __restrict__ row, int x0, int dx, int w, u16 out[8]) {
  union {
    uint4 v;
    u16 h[8];
  } u;
  u.v = *reinterpret_cast<const uint4*>(row + x0);
  if (dx == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = u.h[i];
  } else if (dx < 0) {
    out[0] = x0 > 0 ? row[x0 - 1] : (u16)0;
#pragma unroll
    for (int i = 1; i < 8; ++i) out[i] = u.h[i - 1];
  } else {
#pragma unroll
    for (int i = 0; i < 7; ++i) out[i] = u.h[i + 1];
    out[7] = x0 + 8 < w ? row[x0 + 8] : (u16)0;
  }
}

// the same for an FP32 row, rounded to bf16
__device__ __forceinline__ void row8_f32(const float* __restrict__ row, int x0, int dx, int w, u16 out[8]) {
  const float4 a = *reinterpret_cast<const float4*>(row + x0);
  const float4 b = *reinterpret_cast<const float4*>(row + x0 + 4);
  const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  if (dx == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = bf_bits(v[i]);
  } else if (dx < 0) {
    out[0] = x0 > 0 ? bf_bits(row[x0 - 1]) : (u16)0;
#pragma unroll
    for (int i = 1; i < 8; ++i) out[i] = bf_bits(v[i - 1]);
  } else {
#pragma unroll
    for (int i = 0; i < 7; ++i) out[i] = bf_bits(v[i + 1]);
    out[7] = x0 + 8 < w ? bf_bits(row[x0 + 8]) : (u16)0;
  }
}

// Same tile and MFMA schedule as k_conv_gemm (so the same sums in the same order). a.wt holds
// the weights for FWD and their [C_in][C_out x taps] transpose for BWD_DATA.
template <int MODE, int KS>
__global__ __launch_bounds__(GT) void k_conv_gemm_vec(GemmArgs a) {
  constexpr int KK2 = KS * KS, R = KS / 2;
  __shared__ __align__(16) u16 As[TM * LDP];
  __shared__ __align__(16) u16 Bt[TN * LDP];  // transposed: Bt[n][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kbeg = MODE == GEMM_BWD_W ? blockIdx.z * a.kchunk : 0;
  const int kend = MODE == GEMM_BWD_W ? min(a.K, kbeg + a.kchunk) : a.K;
  const int w = 1 << a.w_sh, wmask = w - 1, hwmask = (1 << a.hw_sh) - 1;
  const View& v = a.x;
  // FWD / BWD_DATA B roles: k row tid / 8 of the step, pixels nb .. nb + 7 of one image row
  const int nb = n0 + ((tid & 7) << 3);
  const bool nb_ok = nb < a.N;
  const int bb = nb >> a.hw_sh, bp = nb & hwmask, by = bp >> a.w_sh, bx0 = bp & wmask;
  v4f acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    // ---- A tile 64 x 32
    if (MODE != GEMM_BWD_W) {
      const u16* wt = reinterpret_cast<const u16*>(a.wt);
      if (a.vec_a) {
        const int mm = tid >> 2, kk = (tid & 3) * 8;
        const int m = m0 + mm, k = k0 + kk;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (m < a.M && k < a.K) val = *reinterpret_cast<const uint4*>(wt + (long)m * a.K + k);
        *reinterpret_cast<uint4*>(&As[mm * LDP + kk]) = val;
      } else {
#pragma unroll
        for (int i = 0; i < TM * TK / GT; ++i) {
          const int e = tid + i * GT;
          const int mm = e / TK, kk = e % TK;
          const int m = m0 + mm, k = k0 + kk;
          As[mm * LDP + kk] = (m < a.M && k < a.K) ? wt[(long)m * a.K + k] : (u16)0;
        }
      }
    } else {  // A(co, (b, p)) = dZ[b][co][p]: 8 consecutive pixels of one image
      const int mm = tid >> 2, kg = (tid & 3) * 8;
      const int m = m0 + mm, k = k0 + kg;
      u16 h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (m < a.M && k < kend) {
        const int b = k >> a.hw_sh, p = k & hwmask;
        row8_f32((const float*)a.y.p + (((long)b * a.y.ctot + m) << a.hw_sh), p, 0, 1 << a.hw_sh, h);
      }
      uint4 val;
      val.x = h[0] | ((unsigned)h[1] << 16);
      val.y = h[2] | ((unsigned)h[3] << 16);
      val.z = h[4] | ((unsigned)h[5] << 16);
      val.w = h[6] | ((unsigned)h[7] << 16);
      *reinterpret_cast<uint4*>(&As[mm * LDP + kg]) = val;
    }
    // ---- B tile 32 x 64 (stored transposed)
    if (MODE == GEMM_FWD || MODE == GEMM_BWD_DATA) {
      const int kk = tid >> 3, k = k0 + kk;
      u16 h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (nb_ok && k < kend) {
        const int ch = k / KK2, d = k - ch * KK2;
        int dy = d / KS - R, dx = d % KS - R;
        if (MODE == GEMM_BWD_DATA) {
          dy = -dy;
          dx = -dx;
        }
        const int sy = by + dy;
        if (sy >= 0 && sy < v.h) {
          const long off = (((long)bb * v.ctot + ch) << a.hw_sh) + ((long)sy << a.w_sh);
          if (MODE == GEMM_FWD) row8_bf16((const u16*)v.p + off, bx0, dx, w, h);
          else row8_f32((const float*)v.p + off, bx0, dx, w, h);
        }
      }
      const int ng = (tid & 7) << 3;
#pragma unroll
      for (int i = 0; i < 8; ++i) Bt[(ng + i) * LDP + kk] = h[i];
    } else {  // B((b, p), (ci, d)) = x[b][ci][p + d]: 8 consecutive pixels of one row
      const int nn = tid >> 2, kg = (tid & 3) * 8;
      const int n = n0 + nn, k = k0 + kg;
      u16 h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (n < a.N && k < kend) {
        const int b = k >> a.hw_sh, p = k & hwmask, yy = p >> a.w_sh, x0 = p & wmask;
        const int ci = n / KK2, d = n - ci * KK2;
        const int sy = yy + d / KS - R, dx = d % KS - R;
        if (sy >= 0 && sy < v.h)
          row8_bf16((const u16*)v.p + (((long)b * v.ctot + ci) << a.hw_sh) + ((long)sy << a.w_sh), x0, dx, w, h);
      }
      uint4 val;
      val.x = h[0] | ((unsigned)h[1] << 16);
      val.y = h[2] | ((unsigned)h[3] << 16);
      val.z = h[4] | ((unsigned)h[5] << 16);
      val.w = h[6] | ((unsigned)h[7] << 16);
      *reinterpret_cast<uint4*>(&Bt[nn * LDP + kg]) = val;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const v8bf va = *reinterpret_cast<const v8bf*>(&As[(wm + i * 16 + (lane & 15)) * LDP + 8 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const v8bf vb = *reinterpret_cast<const v8bf*>(&Bt[(wn + j * 16 + (lane & 15)) * LDP + 8 * (lane >> 4)]);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= a.M || n >= a.N) continue;
        float val = acc[i][j][q];
        if (MODE == GEMM_BWD_W) {
          a.partial[((long)blockIdx.z * a.M + m) * a.N + n] = val;
        } else {
          const View& o = a.y;
          const int b = n >> a.hw_sh, p = n & hwmask;
          const long idx = (((long)b * o.ctot + m) << a.hw_sh) + p;
          if (MODE == GEMM_FWD) {
            val += a.bias[m];
            if (a.relu) val = fmaxf(val, 0.f);
            if (a.out_f32) ((float*)o.p)[idx] = val;
            else ((bf16*)o.p)[idx] = to_bf(val);
          } else {
            ((float*)o.p)[idx] = val;
          }
        }
      }
}

// WT[ci][co * taps + d] = W[co][ci][d] (bf16), the data-gradient GEMM's weight rows
__global__ void k_transpose_w(const bf16* __restrict__ w, bf16* __restrict__ wt, int cout, int cin, int kk2) {
  const long total = (long)cout * cin * kk2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int d = (int)(i % kk2);
    const long r = i / kk2;
    const int ci = (int)(r % cin), co = (int)(r / cin);
    wt[((long)ci * cout + co) * kk2 + d] = w[i];
  }
}

// vec: the vector-operand form (image width >= 8; for BWD_DATA a.wt is the transposed copy)
template <int MODE>
static void launch_gemm(const GemmArgs& a, int ks, dim3 g, hipStream_t st, bool vec) {
  if (vec) {
    if (ks == 3) hipLaunchKernelGGL((k_conv_gemm_vec<MODE, 3>), g, dim3(GT), 0, st, a);
    else hipLaunchKernelGGL((k_conv_gemm_vec<MODE, 1>), g, dim3(GT), 0, st, a);
  } else {
    if (ks == 3) hipLaunchKernelGGL((k_conv_gemm<MODE, 3>), g, dim3(GT), 0, st, a);
    else hipLaunchKernelGGL((k_conv_gemm<MODE, 1>), g, dim3(GT), 0, st, a);
  }
}

static int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

// dW = sum over chunks (fixed order); db[co] = sum over (b, p) of dZ (block partials, fixed order)
__global__ void k_sum_chunks(const float* __restrict__ part, int chunks, long mn, float* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mn) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += part[(long)c * mn + i];
  out[i] = s;
}

// db[co] = sum over (b, p) of dZ: slice sums (grid cout x parts), then one fixed-order sum per
// channel -- deterministic, and enough blocks to fill the GPU at full resolution
__global__ __launch_bounds__(256) void k_bias_partial(View dz, int batch, int hw_sh, long slice,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  const int co = blockIdx.x, parts = gridDim.y;
  const long total = (long)batch << hw_sh, hwm = (1L << hw_sh) - 1;
  const long lo = blockIdx.y * slice, hi = min(total, lo + slice);
  double s = 0.0;
  for (long e = lo + threadIdx.x; e < hi; e += 256)
    s += ((const float*)dz.p)[(((e >> hw_sh) * dz.ctot + co) << hw_sh) + (e & hwm)];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[(long)co * parts + blockIdx.y] = s;
}

__global__ __launch_bounds__(64) void k_bias_final(const double* __restrict__ part, int parts,
                                                   float* __restrict__ db) {
  double s = 0.0;
  for (int i = threadIdx.x; i < parts; i += 64) s += part[(long)blockIdx.x * parts + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) db[blockIdx.x] = (float)s;
}

// Elementwise layers on a (pixel blocks x planes) grid: plane = b * C + c (one block row per
// sample and channel), pixels of the plane strided over blockIdx.x; widths are powers of two.
__device__ __forceinline__ int lg2(int v) { return __ffs(v) - 1; }

// g <- g * (a > 0): gradient through the ReLU (a is the post-activation, bf16)
__global__ void k_relu_grad(View g, View a) {
  const int plane = blockIdx.y, b = plane / g.c, c = plane - b * g.c;
  const int hw = g.h * g.w;
  float* gp = (float*)g.p + ((long)b * g.ctot + c) * hw;
  const bf16* ap = (const bf16*)a.p + ((long)b * a.ctot + c) * hw;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < hw; p += gridDim.x * blockDim.x)
    if (!((float)ap[p] > 0.f)) gp[p] = 0.f;
}

// 2 x 2 max pooling; arg holds the winning offset (0..3) for the backward pass
__global__ void k_pool_fwd(View in, View out, uint8_t* __restrict__ arg) {
  const int plane = blockIdx.y, b = plane / out.c, c = plane - b * out.c;
  const int ohw = out.h * out.w, osh = lg2(out.w);
  const bf16* src = (const bf16*)in.p + ((long)b * in.ctot + c) * in.h * in.w;
  bf16* dst = (bf16*)out.p + ((long)b * out.ctot + c) * ohw;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ohw; p += gridDim.x * blockDim.x) {
    const int y = p >> osh, x = p & (out.w - 1);
    float best = -INFINITY;
    int bi = 0;
    for (int q = 0; q < 4; ++q) {
      const float v = (float)src[(2 * y + (q >> 1)) * in.w + 2 * x + (q & 1)];
      if (v > best) {
        best = v;
        bi = q;
      }
    }
    dst[p] = to_bf(best);
    arg[(long)plane * ohw + p] = (uint8_t)bi;
  }
}

// gin (f32, the skip gradient already there) += scatter of gout to the winning offsets
__global__ void k_pool_bwd(View gout, View gin, const uint8_t* __restrict__ arg) {
  const int plane = blockIdx.y, b = plane / gout.c, c = plane - b * gout.c;
  const int ohw = gout.h * gout.w, osh = lg2(gout.w);
  const float* src = (const float*)gout.p + ((long)b * gout.ctot + c) * ohw;
  float* dst = (float*)gin.p + ((long)b * gin.ctot + c) * gin.h * gin.w;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ohw; p += gridDim.x * blockDim.x) {
    const int y = p >> osh, x = p & (gout.w - 1);
    const int q = arg[(long)plane * ohw + p];
    dst[(2 * y + (q >> 1)) * gin.w + 2 * x + (q & 1)] += src[p];
  }
}

// nearest x2 upsampling (bf16) of in's channels into out's first channels, and its gradient
// (sum of the four children)
__global__ void k_up_fwd(View in, View out) {
  const int plane = blockIdx.y, b = plane / in.c, c = plane - b * in.c;
  const int ohw = out.h * out.w, osh = lg2(out.w);
  const bf16* src = (const bf16*)in.p + ((long)b * in.ctot + c) * in.h * in.w;
  bf16* dst = (bf16*)out.p + ((long)b * out.ctot + c) * ohw;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ohw; p += gridDim.x * blockDim.x) {
    const int y = p >> osh, x = p & (out.w - 1);
    dst[p] = src[(y >> 1) * in.w + (x >> 1)];
  }
}
__global__ void k_up_bwd(View gout, View gin) {
  const int plane = blockIdx.y, b = plane / gin.c, c = plane - b * gin.c;
  const int ihw = gin.h * gin.w, ish = lg2(gin.w);
  const float* src = (const float*)gout.p + ((long)b * gout.ctot + c) * gout.h * gout.w;
  float* dst = (float*)gin.p + ((long)b * gin.ctot + c) * ihw;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ihw; p += gridDim.x * blockDim.x) {
    const int y = p >> ish, x = p & (gin.w - 1);
    dst[p] = src[(2 * y) * gout.w + 2 * x] + src[(2 * y) * gout.w + 2 * x + 1] + src[(2 * y + 1) * gout.w + 2 * x] +
             src[(2 * y + 1) * gout.w + 2 * x + 1];
  }
}

// weighted mean squared error: sum w(p) (y - t)^2 over all outputs, w an s x s map shared by the
// channels and samples (NULL: 1); g = 2 w (y - t) / norm with norm = batch * 2 * sum_p w(p);
// per-block FP64 partial sums of w (y - t)^2
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const bf16* __restrict__ t, long count,
                                             const float* __restrict__ w, long ss, double norm,
                                             float* __restrict__ g, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  const float sc = (float)(2.0 / norm);
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long)gridDim.x * 256) {
    const float d = y[i] - (float)t[i];
    const float wi = w ? w[i % ss] : 1.f;
    s += (double)wi * d * d;
    if (g) g[i] = sc * wi * d;
  }
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Adam on FP32 master weights; refreshes the BF16 copy
__global__ void k_adam(float* __restrict__ w, float* __restrict__ mo, float* __restrict__ ve,
                       const float* __restrict__ g, bf16* __restrict__ wb, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * mo[i] + (1.f - b1) * gi;
  const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
  mo[i] = m;
  ve[i] = v;
  const float wn = w[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  w[i] = wn;
  wb[i] = to_bf(wn);
}

__global__ void k_to_bf(const float* __restrict__ s, bf16* __restrict__ d, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = to_bf(s[i]);
}

// (pixel blocks, planes) grid of an elementwise layer over `batch` samples of view v
inline dim3 plane_grid(const View& v, int batch) {
  const long hw = (long)v.h * v.w;
  return dim3((unsigned)std::min<long>((hw + 255) / 256, 64), (unsigned)(batch * v.c));
}

inline dim3 grid_of(long n) {
  long b = (n + 255) / 256;
  return dim3((unsigned)std::min<long>(b, 65535L * 4));
}

}  // namespace

// ---------------------------------------------------------------- host side
struct NnLayer {
  int cin, cout, ks, relu;
  long woff, boff;  // offsets in the parameter vector
};

struct NnNet {
  int s = 256, c = 32, batch = 8;
  std::vector<NnLayer> layers;
  long nparam = 0;
  float *w = nullptr, *mo = nullptr, *ve = nullptr, *g = nullptr;
  bf16* wb = nullptr;
  bf16* wbt = nullptr;  // per layer [C_in][C_out x taps] transpose of wb (data-gradient weights)
  double* bpart = nullptr;  // bias-gradient slice sums, 8c x BIAS_PARTS
  float* partial = nullptr;
  long partial_cap = 0;
  double* lpart = nullptr;
  long step = 0;
  // activations (bf16) and their gradients (f32)
  struct Buf {
    int c, h;
    bf16* a = nullptr;
    float* g = nullptr;
  };
  std::vector<Buf> bufs;
  uint8_t* arg[3] = {nullptr, nullptr, nullptr};
  float* out = nullptr;   // batch x 2 x s x s
  float* gout = nullptr;
  float* lw = nullptr;    // s x s loss weight (NULL: uniform)
  double lw_sum = 0.0;    // sum of the weight (s^2 when uniform)
  bf16* target = nullptr; // batch x 2 x s x s (staging)
  hipStream_t stream = nullptr;
};

// Buffer indices (see build): 0 x, 1 a0, 2 cat0, 3 p0, 4 a1, 5 cat1, 6 p1, 7 a2, 8 cat2, 9 p2,
// 10 a3, 11 bneck, 12 a4, 13 d2, 14 a5, 15 d1, 16 a6, 17 d0
enum { B_X, B_A0, B_CAT0, B_P0, B_A1, B_CAT1, B_P1, B_A2, B_CAT2, B_P2, B_A3, B_BN, B_A4, B_D2, B_A5, B_D1, B_A6,
       B_D0, B_COUNT };

static View vw(const NnNet& n, int bi, int c0, int c, bool grad) {
  const auto& b = n.bufs[bi];
  const long hw = (long)b.h * b.h;
  void* base = grad ? (void*)(b.g + (long)c0 * hw) : (void*)(b.a + (long)c0 * hw);
  return View{base, c, b.c, b.h, b.h};
}

int nn_layer_table(int c, std::vector<NnLayer>* out, long* nparam) {
  const int tab[15][4] = {{1, c, 3, 1},          {c, c, 3, 1},          {c, 2 * c, 3, 1},     {2 * c, 2 * c, 3, 1},
                          {2 * c, 4 * c, 3, 1},  {4 * c, 4 * c, 3, 1},  {4 * c, 8 * c, 3, 1}, {8 * c, 8 * c, 3, 1},
                          {12 * c, 4 * c, 3, 1}, {4 * c, 4 * c, 3, 1},  {6 * c, 2 * c, 3, 1}, {2 * c, 2 * c, 3, 1},
                          {3 * c, c, 3, 1},      {c, c, 3, 1},          {c, 2, 1, 0}};
  out->clear();
  long off = 0;
  for (auto& t : tab) {
    NnLayer l{t[0], t[1], t[2], t[3], 0, 0};
    l.woff = off;
    off += (long)l.cout * l.cin * l.ks * l.ks;
    l.boff = off;
    off += l.cout;
    out->push_back(l);
  }
  *nparam = off;
  return 15;
}

static hipError_t conv_fwd(NnNet& n, int li, View x, View y, bool out_f32) {
  const NnLayer& L = n.layers[li];
  GemmArgs a{};
  a.x = x;
  a.y = y;
  a.wt = n.wb + L.woff;
  a.bias = n.w + L.boff;
  a.batch = n.batch;
  a.cin = L.cin;
  a.cout = L.cout;
  a.relu = L.relu;
  a.out_f32 = out_f32 ? 1 : 0;
  a.M = L.cout;
  a.N = n.batch * x.h * x.w;
  a.K = L.cin * L.ks * L.ks;
  a.w_sh = ilog2(x.w);
  a.hw_sh = ilog2(x.h * x.w);
  a.vec_a = (a.K % 8 == 0 && L.woff % 8 == 0 && ((uintptr_t)n.wb & 15) == 0) ? 1 : 0;
  dim3 g((a.N + TN - 1) / TN, (a.M + TM - 1) / TM, 1);
  launch_gemm<GEMM_FWD>(a, L.ks, g, n.stream, x.w >= 8);
  return hipGetLastError();
}

// dz: gradient w.r.t. the layer's pre-activation output (f32 view); x: its input (bf16 view);
// dx (nullable): gradient w.r.t. the input, overwritten.
static hipError_t conv_bwd(NnNet& n, int li, View x, View dz, View* dx) {
  const NnLayer& L = n.layers[li];
  const int kk2 = L.ks * L.ks;
  {  // weight gradient, split over pixels
    GemmArgs a{};
    a.x = x;
    a.y = dz;
    a.batch = n.batch;
    a.cin = L.cin;
    a.cout = L.cout;
    a.M = L.cout;
    a.N = L.cin * kk2;
    a.K = n.batch * x.h * x.w;
    a.w_sh = ilog2(x.w);
    a.hw_sh = ilog2(x.h * x.w);
    int chunks = std::max(1, std::min(256, a.K / 4096));
    a.kchunk = ((a.K + chunks - 1) / chunks + TK - 1) / TK * TK;
    chunks = (a.K + a.kchunk - 1) / a.kchunk;
    if ((long)chunks * a.M * a.N > n.partial_cap) return hipErrorInvalidValue;
    a.partial = n.partial;
    dim3 g((a.N + TN - 1) / TN, (a.M + TM - 1) / TM, chunks);
    launch_gemm<GEMM_BWD_W>(a, L.ks, g, n.stream, x.w >= 8);
    const long mn = (long)a.M * a.N;
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, n.stream, n.partial, chunks,
                       mn, n.g + L.woff);
    {
      const int hw_sh = ilog2(dz.h * dz.w);
      const long total = (long)n.batch << hw_sh;
      const int parts = (int)std::max<long>(1, std::min<long>(BIAS_PARTS, total / 4096));
      const long slice = (total + parts - 1) / parts;
      hipLaunchKernelGGL(k_bias_partial, dim3(L.cout, parts), dim3(256), 0, n.stream, dz, n.batch, hw_sh, slice,
                         n.bpart);
      hipLaunchKernelGGL(k_bias_final, dim3(L.cout), dim3(64), 0, n.stream, n.bpart, parts, n.g + L.boff);
    }
  }
  if (dx) {
    GemmArgs a{};
    a.x = dz;
    a.y = *dx;
    const bool vec = dz.w >= 8;
    a.wt = (vec ? n.wbt : n.wb) + L.woff;
    a.batch = n.batch;
    a.cin = L.cin;
    a.cout = L.cout;
    a.M = L.cin;
    a.N = n.batch * dz.h * dz.w;
    a.K = L.cout * kk2;
    a.w_sh = ilog2(dz.w);
    a.hw_sh = ilog2(dz.h * dz.w);
    a.vec_a = (a.K % 8 == 0 && L.woff % 8 == 0 && ((uintptr_t)n.wbt & 15) == 0) ? 1 : 0;
    dim3 g((a.N + TN - 1) / TN, (a.M + TM - 1) / TM, 1);
    launch_gemm<GEMM_BWD_DATA>(a, L.ks, g, n.stream, vec);
  }
  return hipGetLastError();
}


hipError_t nn_forward(NnNet& n) {
  const int c = n.c;
  hipError_t e;
  auto fwd = [&](int li, View x, View y) { return conv_fwd(n, li, x, y, false); };
  if ((e = fwd(0, vw(n, B_X, 0, 1, false), vw(n, B_A0, 0, c, false)))) return e;
  if ((e = fwd(1, vw(n, B_A0, 0, c, false), vw(n, B_CAT0, 2 * c, c, false)))) return e;
  View p0 = vw(n, B_P0, 0, c, false);
  hipLaunchKernelGGL(k_pool_fwd, plane_grid(p0, n.batch), dim3(256), 0, n.stream, vw(n, B_CAT0, 2 * c, c, false), p0,
                     n.arg[0]);
  if ((e = fwd(2, p0, vw(n, B_A1, 0, 2 * c, false)))) return e;
  if ((e = fwd(3, vw(n, B_A1, 0, 2 * c, false), vw(n, B_CAT1, 4 * c, 2 * c, false)))) return e;
  View p1 = vw(n, B_P1, 0, 2 * c, false);
  hipLaunchKernelGGL(k_pool_fwd, plane_grid(p1, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_CAT1, 4 * c, 2 * c, false), p1, n.arg[1]);
  if ((e = fwd(4, p1, vw(n, B_A2, 0, 4 * c, false)))) return e;
  if ((e = fwd(5, vw(n, B_A2, 0, 4 * c, false), vw(n, B_CAT2, 8 * c, 4 * c, false)))) return e;
  View p2 = vw(n, B_P2, 0, 4 * c, false);
  hipLaunchKernelGGL(k_pool_fwd, plane_grid(p2, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_CAT2, 8 * c, 4 * c, false), p2, n.arg[2]);
  if ((e = fwd(6, p2, vw(n, B_A3, 0, 8 * c, false)))) return e;
  if ((e = fwd(7, vw(n, B_A3, 0, 8 * c, false), vw(n, B_BN, 0, 8 * c, false)))) return e;
  View up2 = vw(n, B_CAT2, 0, 8 * c, false);
  hipLaunchKernelGGL(k_up_fwd, plane_grid(up2, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_BN, 0, 8 * c, false), up2);
  if ((e = fwd(8, vw(n, B_CAT2, 0, 12 * c, false), vw(n, B_A4, 0, 4 * c, false)))) return e;
  if ((e = fwd(9, vw(n, B_A4, 0, 4 * c, false), vw(n, B_D2, 0, 4 * c, false)))) return e;
  View up1 = vw(n, B_CAT1, 0, 4 * c, false);
  hipLaunchKernelGGL(k_up_fwd, plane_grid(up1, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_D2, 0, 4 * c, false), up1);
  if ((e = fwd(10, vw(n, B_CAT1, 0, 6 * c, false), vw(n, B_A5, 0, 2 * c, false)))) return e;
  if ((e = fwd(11, vw(n, B_A5, 0, 2 * c, false), vw(n, B_D1, 0, 2 * c, false)))) return e;
  View up0 = vw(n, B_CAT0, 0, 2 * c, false);
  hipLaunchKernelGGL(k_up_fwd, plane_grid(up0, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_D1, 0, 2 * c, false), up0);
  if ((e = fwd(12, vw(n, B_CAT0, 0, 3 * c, false), vw(n, B_A6, 0, c, false)))) return e;
  if ((e = fwd(13, vw(n, B_A6, 0, c, false), vw(n, B_D0, 0, c, false)))) return e;
  View out{n.out, 2, 2, n.s, n.s};
  if ((e = conv_fwd(n, 14, vw(n, B_D0, 0, c, false), out, true))) return e;
  return hipGetLastError();
}

// loss over the batch (returned via lpart partials, count = batch*2*s*s); with_grad: backward
hipError_t nn_loss(NnNet& n, bool with_grad, int* nparts) {
  const long count = (long)n.batch * 2 * n.s * n.s;
  const int nb = (int)std::min<long>((count + 255) / 256, 1024);
  const double norm = (double)n.batch * 2 * n.lw_sum;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, n.stream, n.out, n.target, count, n.lw, (long)n.s * n.s, norm,
                     with_grad ? n.gout : nullptr, n.lpart);
  *nparts = nb;
  return hipGetLastError();
}

hipError_t nn_backward(NnNet& n) {
  const int c = n.c;
  hipError_t e;
  for (const NnLayer& L : n.layers) {
    const long cnt = (long)L.cout * L.cin * L.ks * L.ks;
    hipLaunchKernelGGL(k_transpose_w, grid_of(cnt), dim3(256), 0, n.stream, n.wb + L.woff, n.wbt + L.woff, L.cout,
                       L.cin, L.ks * L.ks);
  }
  auto relu_g = [&](View g, View a) {
    hipLaunchKernelGGL(k_relu_grad, plane_grid(g, n.batch), dim3(256), 0, n.stream, g, a);
  };
  View gout{n.gout, 2, 2, n.s, n.s};
  View gd0 = vw(n, B_D0, 0, c, true);
  if ((e = conv_bwd(n, 14, vw(n, B_D0, 0, c, false), gout, &gd0))) return e;
  // level 0 decoder
  relu_g(gd0, vw(n, B_D0, 0, c, false));
  View ga6 = vw(n, B_A6, 0, c, true);
  if ((e = conv_bwd(n, 13, vw(n, B_A6, 0, c, false), gd0, &ga6))) return e;
  relu_g(ga6, vw(n, B_A6, 0, c, false));
  View gcat0 = vw(n, B_CAT0, 0, 3 * c, true);
  if ((e = conv_bwd(n, 12, vw(n, B_CAT0, 0, 3 * c, false), ga6, &gcat0))) return e;
  View gd1 = vw(n, B_D1, 0, 2 * c, true);
  hipLaunchKernelGGL(k_up_bwd, plane_grid(gd1, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_CAT0, 0, 2 * c, true), gd1);
  // level 1 decoder
  relu_g(gd1, vw(n, B_D1, 0, 2 * c, false));
  View ga5 = vw(n, B_A5, 0, 2 * c, true);
  if ((e = conv_bwd(n, 11, vw(n, B_A5, 0, 2 * c, false), gd1, &ga5))) return e;
  relu_g(ga5, vw(n, B_A5, 0, 2 * c, false));
  View gcat1 = vw(n, B_CAT1, 0, 6 * c, true);
  if ((e = conv_bwd(n, 10, vw(n, B_CAT1, 0, 6 * c, false), ga5, &gcat1))) return e;
  View gd2 = vw(n, B_D2, 0, 4 * c, true);
  hipLaunchKernelGGL(k_up_bwd, plane_grid(gd2, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_CAT1, 0, 4 * c, true), gd2);
  // level 2 decoder
  relu_g(gd2, vw(n, B_D2, 0, 4 * c, false));
  View ga4 = vw(n, B_A4, 0, 4 * c, true);
  if ((e = conv_bwd(n, 9, vw(n, B_A4, 0, 4 * c, false), gd2, &ga4))) return e;
  relu_g(ga4, vw(n, B_A4, 0, 4 * c, false));
  View gcat2 = vw(n, B_CAT2, 0, 12 * c, true);
  if ((e = conv_bwd(n, 8, vw(n, B_CAT2, 0, 12 * c, false), ga4, &gcat2))) return e;
  View gbn = vw(n, B_BN, 0, 8 * c, true);
  hipLaunchKernelGGL(k_up_bwd, plane_grid(gbn, n.batch), dim3(256), 0, n.stream,
                     vw(n, B_CAT2, 0, 8 * c, true), gbn);
  // bottleneck
  relu_g(gbn, vw(n, B_BN, 0, 8 * c, false));
  View ga3 = vw(n, B_A3, 0, 8 * c, true);
  if ((e = conv_bwd(n, 7, vw(n, B_A3, 0, 8 * c, false), gbn, &ga3))) return e;
  relu_g(ga3, vw(n, B_A3, 0, 8 * c, false));
  View gp2 = vw(n, B_P2, 0, 4 * c, true);
  if ((e = conv_bwd(n, 6, vw(n, B_P2, 0, 4 * c, false), ga3, &gp2))) return e;
  View ge2 = vw(n, B_CAT2, 8 * c, 4 * c, true);
  hipLaunchKernelGGL(k_pool_bwd, plane_grid(gp2, n.batch), dim3(256), 0, n.stream, gp2, ge2, n.arg[2]);
  // level 2 encoder
  relu_g(ge2, vw(n, B_CAT2, 8 * c, 4 * c, false));
  View ga2 = vw(n, B_A2, 0, 4 * c, true);
  if ((e = conv_bwd(n, 5, vw(n, B_A2, 0, 4 * c, false), ge2, &ga2))) return e;
  relu_g(ga2, vw(n, B_A2, 0, 4 * c, false));
  View gp1 = vw(n, B_P1, 0, 2 * c, true);
  if ((e = conv_bwd(n, 4, vw(n, B_P1, 0, 2 * c, false), ga2, &gp1))) return e;
  View ge1 = vw(n, B_CAT1, 4 * c, 2 * c, true);
  hipLaunchKernelGGL(k_pool_bwd, plane_grid(gp1, n.batch), dim3(256), 0, n.stream, gp1, ge1, n.arg[1]);
  // level 1 encoder
  relu_g(ge1, vw(n, B_CAT1, 4 * c, 2 * c, false));
  View ga1 = vw(n, B_A1, 0, 2 * c, true);
  if ((e = conv_bwd(n, 3, vw(n, B_A1, 0, 2 * c, false), ge1, &ga1))) return e;
  relu_g(ga1, vw(n, B_A1, 0, 2 * c, false));
  View gp0 = vw(n, B_P0, 0, c, true);
  if ((e = conv_bwd(n, 2, vw(n, B_P0, 0, c, false), ga1, &gp0))) return e;
  View ge0 = vw(n, B_CAT0, 2 * c, c, true);
  hipLaunchKernelGGL(k_pool_bwd, plane_grid(gp0, n.batch), dim3(256), 0, n.stream, gp0, ge0, n.arg[0]);
  // level 0 encoder
  relu_g(ge0, vw(n, B_CAT0, 2 * c, c, false));
  View ga0 = vw(n, B_A0, 0, c, true);
  if ((e = conv_bwd(n, 1, vw(n, B_A0, 0, c, false), ge0, &ga0))) return e;
  relu_g(ga0, vw(n, B_A0, 0, c, false));
  if ((e = conv_bwd(n, 0, vw(n, B_X, 0, 1, false), ga0, nullptr))) return e;
  return hipGetLastError();
}

hipError_t nn_adam(NnNet& n, float lr) {
  n.step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = (float)(1.0 - std::pow(0.9, (double)n.step)), c2 = (float)(1.0 - std::pow(0.999, (double)n.step));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n.nparam + 255) / 256)), dim3(256), 0, n.stream, n.w, n.mo, n.ve, n.g,
                     n.wb, n.nparam, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t nn_refresh_bf16(NnNet& n) {
  hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((n.nparam + 255) / 256)), dim3(256), 0, n.stream, n.w, n.wb, n.nparam);
  return hipGetLastError();
}

static hipError_t nn_alloc1(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e == hipSuccess) e = hipMemset(*p, 0, bytes ? bytes : 8);
  return e;
}

hipError_t nn_create(NnNet& n, int s, int c, int batch) {
  n.s = s;
  n.c = c;
  n.batch = batch;
  n.lw_sum = (double)s * s;
  nn_layer_table(c, &n.layers, &n.nparam);
  const int ch[B_COUNT] = {1, c, 3 * c, c, 2 * c, 6 * c, 2 * c, 4 * c, 12 * c, 4 * c, 8 * c, 8 * c, 4 * c, 4 * c,
                           2 * c, 2 * c, c, c};
  const int lv[B_COUNT] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0};
  n.bufs.resize(B_COUNT);
  hipError_t e = hipSuccess;
  for (int i = 0; i < B_COUNT && e == hipSuccess; ++i) {
    n.bufs[i].c = ch[i];
    n.bufs[i].h = s >> lv[i];
    const size_t cnt = (size_t)batch * ch[i] * n.bufs[i].h * n.bufs[i].h;
    e = nn_alloc1((void**)&n.bufs[i].a, cnt * sizeof(bf16));
    if (e == hipSuccess) e = nn_alloc1((void**)&n.bufs[i].g, cnt * sizeof(float));
  }
  for (int l = 0; l < 3 && e == hipSuccess; ++l) {
    const size_t cnt = (size_t)batch * (c << l) * (s >> (l + 1)) * (s >> (l + 1));
    e = nn_alloc1((void**)&n.arg[l], cnt);
  }
  const size_t outc = (size_t)batch * 2 * s * s;
  if (e == hipSuccess) e = nn_alloc1((void**)&n.out, outc * sizeof(float));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.gout, outc * sizeof(float));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.target, outc * sizeof(bf16));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.w, n.nparam * sizeof(float));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.mo, n.nparam * sizeof(float));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.ve, n.nparam * sizeof(float));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.g, n.nparam * sizeof(float));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.wb, n.nparam * sizeof(bf16));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.wbt, n.nparam * sizeof(bf16));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.bpart, (size_t)8 * c * BIAS_PARTS * sizeof(double));
  if (e == hipSuccess) e = nn_alloc1((void**)&n.lpart, 1024 * sizeof(double));
  // weight-gradient partial tiles: the largest chunks x Cout x Cin ks^2 over the layers
  static const int level[15] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
  long cap = 0;
  for (size_t li = 0; li < n.layers.size(); ++li) {
    const NnLayer& L = n.layers[li];
    const int hh = s >> level[li];
    const long K = (long)batch * hh * hh;
    const int chunks = std::max(1, std::min(256, (int)(K / 4096)));
    const long kc = ((K + chunks - 1) / chunks + TK - 1) / TK * TK;
    const long nch = (K + kc - 1) / kc;
    cap = std::max(cap, nch * L.cout * L.cin * L.ks * L.ks);
  }
  n.partial_cap = cap;
  if (e == hipSuccess) e = nn_alloc1((void**)&n.partial, cap * sizeof(float));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return e;
}

void nn_free(NnNet& n) {
  for (auto& b : n.bufs) {
    if (b.a) hipFree(b.a);
    if (b.g) hipFree(b.g);
  }
  n.bufs.clear();
  for (auto& p : n.arg)
    if (p) hipFree(p);
  for (void* p : {(void*)n.out, (void*)n.gout, (void*)n.target, (void*)n.w, (void*)n.mo, (void*)n.ve, (void*)n.g,
                  (void*)n.wb, (void*)n.wbt, (void*)n.bpart, (void*)n.lpart, (void*)n.partial, (void*)n.lw})
    if (p) hipFree(p);
}


__global__ void k_gather_rows(const bf16* __restrict__ src, const int64_t* __restrict__ idx, long row, int count,
                              bf16* __restrict__ dst) {
  const long total = (long)count * row;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / row, q = i - r * row;
    dst[i] = src[idx[r] * row + q];
  }
}

// Network input from raw amplitudes, one block per sample: x = log(1 + max(A, 0)) / max over the
// sample (a sample whose maximum is 0 maps to 0).
__global__ __launch_bounds__(256) void k_prep_amplitudes(const float* __restrict__ amp, long ss,
                                                         bf16* __restrict__ x) {
  __shared__ float red[4];
  const float* a = amp + blockIdx.x * ss;
  bf16* o = x + blockIdx.x * ss;
  float mx = 0.f;
  for (long i = threadIdx.x; i < ss; i += 256) mx = fmaxf(mx, log1pf(fmaxf(a[i], 0.f)));
  mx = block_max<256>(mx, red);
  const float inv = mx > 0.f ? 1.f / mx : 0.f;
  for (long i = threadIdx.x; i < ss; i += 256) o[i] = to_bf(log1pf(fmaxf(a[i], 0.f)) * inv);
}

}  // namespace ptyk

// ---------------------------------------------------------------- C ABI
using namespace ptyk;

struct ptyk_net {
  NnNet n;
  int device = 0;
  std::string err;
  bf16* data_x = nullptr;
  bf16* data_y = nullptr;
  int64_t ndata = 0;     // samples [0, ndata) hold data
  int64_t data_cap = 0;  // samples the device buffers hold room for
  int64_t* idx = nullptr;
  int64_t idx_cap = 0;
};

static thread_local std::string g_net_err;

#define NCHK(h, call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);                    \
      return e_ == hipErrorOutOfMemory ? PTYK_ENOMEM : PTYK_EDEVICE;                   \
    }                                                                                  \
  } while (0)

extern "C" {

int ptyk_net_create(int32_t device, int32_t size, int32_t base_channels, int32_t batch, uint64_t seed,
                    ptyk_net** out) {
  if (!out || size < 8 || (size & (size - 1)) || base_channels < 1 || batch < 1) {
    g_net_err = "size must be a power of two >= 8; channels and batch >= 1";
    return PTYK_EINVAL;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    g_net_err = "no such HIP device";
    return PTYK_EDEVICE;
  }
  hipSetDevice(device);
  ptyk_net* h = new ptyk_net();
  h->device = device;
  hipError_t e = hipStreamCreateWithFlags(&h->n.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = nn_create(h->n, size, base_channels, batch);
  if (e != hipSuccess) {
    g_net_err = std::string("network allocation: ") + hipGetErrorString(e);
    nn_free(h->n);
    if (h->n.stream) hipStreamDestroy(h->n.stream);
    delete h;
    return e == hipErrorOutOfMemory ? PTYK_ENOMEM : PTYK_EDEVICE;
  }
  // He-normal weights (seeded), zero biases
  std::vector<float> w(h->n.nparam, 0.f);
  std::mt19937_64 g(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (const auto& L : h->n.layers) {
    const double sd = std::sqrt(2.0 / (L.cin * L.ks * L.ks));
    for (long i = 0; i < (long)L.cout * L.cin * L.ks * L.ks; ++i) w[L.woff + i] = (float)(sd * nd(g));
  }
  int r = ptyk_net_set_params(h, w.data());
  if (r) {
    g_net_err = h->err;
    ptyk_net_destroy(h);
    return r;
  }
  *out = h;
  return PTYK_OK;
}

void ptyk_net_destroy(ptyk_net* h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->n.stream) hipStreamSynchronize(h->n.stream);
  nn_free(h->n);
  for (void* p : {(void*)h->data_x, (void*)h->data_y, (void*)h->idx})
    if (p) hipFree(p);
  if (h->n.stream) hipStreamDestroy(h->n.stream);
  delete h;
}

const char* ptyk_net_last_error(const ptyk_net* h) { return h ? h->err.c_str() : g_net_err.c_str(); }

int64_t ptyk_net_num_params(const ptyk_net* h) { return h ? h->n.nparam : -1; }

int ptyk_net_get_params(ptyk_net* h, float* out) {
  if (!h || !out) return PTYK_EINVAL;
  hipSetDevice(h->device);
  NCHK(h, hipMemcpyAsync(out, h->n.w, h->n.nparam * sizeof(float), hipMemcpyDeviceToHost, h->n.stream));
  NCHK(h, hipStreamSynchronize(h->n.stream));
  return PTYK_OK;
}

int ptyk_net_set_params(ptyk_net* h, const float* in) {
  if (!h || !in) return PTYK_EINVAL;
  hipSetDevice(h->device);
  NCHK(h, hipMemcpyAsync(h->n.w, in, h->n.nparam * sizeof(float), hipMemcpyHostToDevice, h->n.stream));
  NCHK(h, nn_refresh_bf16(h->n));
  NCHK(h, hipMemsetAsync(h->n.mo, 0, h->n.nparam * sizeof(float), h->n.stream));
  NCHK(h, hipMemsetAsync(h->n.ve, 0, h->n.nparam * sizeof(float), h->n.stream));
  NCHK(h, hipStreamSynchronize(h->n.stream));
  h->n.step = 0;
  return PTYK_OK;
}

// Per-pixel loss weight (s x s, non-negative, not all zero), shared by both output channels and
// every sample; NULL restores the uniform weight.
int ptyk_net_set_loss_weight(ptyk_net* h, const float* weight) {
  if (!h) return PTYK_EINVAL;
  hipSetDevice(h->device);
  const long ss = (long)h->n.s * h->n.s;
  if (!weight) {
    NCHK(h, hipStreamSynchronize(h->n.stream));
    if (h->n.lw) hipFree(h->n.lw);
    h->n.lw = nullptr;
    h->n.lw_sum = (double)ss;
    return PTYK_OK;
  }
  double sum = 0.0;
  for (long i = 0; i < ss; ++i) {
    if (!(weight[i] >= 0.f) || !std::isfinite(weight[i])) {
      h->err = "loss weights must be finite and non-negative";
      return PTYK_EINVAL;
    }
    sum += weight[i];
  }
  if (!(sum > 0.0)) {
    h->err = "loss weights are all zero";
    return PTYK_EINVAL;
  }
  if (!h->n.lw) NCHK(h, hipMalloc((void**)&h->n.lw, ss * sizeof(float)));
  NCHK(h, hipMemcpyAsync(h->n.lw, weight, ss * sizeof(float), hipMemcpyHostToDevice, h->n.stream));
  NCHK(h, hipStreamSynchronize(h->n.stream));
  h->n.lw_sum = sum;
  return PTYK_OK;
}

// Device room for n samples (BF16); previous samples are dropped.
int ptyk_net_reserve_data(ptyk_net* h, int64_t n) {
  if (!h || n <= 0) return PTYK_EINVAL;
  hipSetDevice(h->device);
  const long ss = (long)h->n.s * h->n.s;
  NCHK(h, hipStreamSynchronize(h->n.stream));
  if (h->data_x) hipFree(h->data_x);
  if (h->data_y) hipFree(h->data_y);
  h->data_x = h->data_y = nullptr;
  h->ndata = h->data_cap = 0;
  NCHK(h, hipMalloc((void**)&h->data_x, (size_t)n * ss * sizeof(bf16)));
  NCHK(h, hipMalloc((void**)&h->data_y, (size_t)n * 2 * ss * sizeof(bf16)));
  h->data_cap = n;
  return PTYK_OK;
}

// BF16 device copy of samples [offset, offset + count): inputs count x s x s, targets
// count x 2 x s x s (float, host); samples up to offset + count become usable.
int ptyk_net_put_data(ptyk_net* h, int64_t offset, const float* inputs, const float* targets, int64_t count) {
  if (!h || !inputs || !targets || offset < 0 || count <= 0) return PTYK_EINVAL;
  if (offset + count > h->data_cap) {
    h->err = "samples beyond the reserved room";
    return PTYK_EINVAL;
  }
  hipSetDevice(h->device);
  const long ss = (long)h->n.s * h->n.s;
  const int64_t slice = 256;
  float* stage = nullptr;
  NCHK(h, hipMalloc((void**)&stage, (size_t)std::min(slice, count) * 2 * ss * sizeof(float)));
  hipError_t e = hipSuccess;
  for (int64_t b = 0; b < count && e == hipSuccess; b += slice) {
    const int64_t cnt = std::min(slice, count - b);
    bf16* dx = h->data_x + (offset + b) * ss;
    bf16* dy = h->data_y + (offset + b) * 2 * ss;
    e = hipMemcpyAsync(stage, inputs + b * ss, (size_t)cnt * ss * sizeof(float), hipMemcpyHostToDevice, h->n.stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((cnt * ss + 255) / 256)), dim3(256), 0, h->n.stream, stage, dx,
                         (long)(cnt * ss));
      e = hipMemcpyAsync(stage, targets + b * 2 * ss, (size_t)cnt * 2 * ss * sizeof(float), hipMemcpyHostToDevice,
                         h->n.stream);
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((cnt * 2 * ss + 255) / 256)), dim3(256), 0, h->n.stream, stage, dy,
                         (long)(cnt * 2 * ss));
      e = hipStreamSynchronize(h->n.stream);  // the staging buffer is refilled next slice
    }
  }
  hipFree(stage);
  NCHK(h, e);
  h->ndata = std::max(h->ndata, offset + count);
  return PTYK_OK;
}

int ptyk_net_set_data(ptyk_net* h, const float* inputs, const float* targets, int64_t n) {
  if (!h || !inputs || !targets || n <= 0) return PTYK_EINVAL;
  int r = ptyk_net_reserve_data(h, n);
  return r ? r : ptyk_net_put_data(h, 0, inputs, targets, n);
}

static int net_load_batch(ptyk_net* h, const int64_t* order, int64_t off) {
  const long ss = (long)h->n.s * h->n.s;
  const int B = h->n.batch;
  if (h->idx_cap < B) {
    if (h->idx) hipFree(h->idx);
    NCHK(h, hipMalloc((void**)&h->idx, B * sizeof(int64_t)));
    h->idx_cap = B;
  }
  NCHK(h, hipMemcpyAsync(h->idx, order + off, B * sizeof(int64_t), hipMemcpyHostToDevice, h->n.stream));
  hipLaunchKernelGGL(k_gather_rows, grid_of((long)B * ss), dim3(256), 0, h->n.stream, h->data_x, h->idx, ss, B,
                     h->n.bufs[B_X].a);
  hipLaunchKernelGGL(k_gather_rows, grid_of((long)B * 2 * ss), dim3(256), 0, h->n.stream, h->data_y, h->idx, 2 * ss,
                     B, h->n.target);
  NCHK(h, hipStreamSynchronize(h->n.stream));  // the order slice is host memory of the caller
  return PTYK_OK;
}

static int net_loss_value(ptyk_net* h, int nparts, double* loss) {
  std::vector<double> p(nparts);
  NCHK(h, hipMemcpyAsync(p.data(), h->n.lpart, nparts * sizeof(double), hipMemcpyDeviceToHost, h->n.stream));
  NCHK(h, hipStreamSynchronize(h->n.stream));
  double s = 0;
  for (double v : p) s += v;
  *loss = s / ((double)h->n.batch * 2 * h->n.lw_sum);
  return PTYK_OK;
}

// Adam steps over the stored samples order[0..count) in batches (count a multiple of batch).
int ptyk_net_train(ptyk_net* h, const int64_t* order, int64_t count, float lr, double* mean_loss) {
  if (!h || !order || count <= 0 || count % h->n.batch) {
    if (h) h->err = "count must be a positive multiple of the batch";
    return PTYK_EINVAL;
  }
  for (int64_t i = 0; i < count; ++i)
    if (order[i] < 0 || order[i] >= h->ndata) {
      h->err = "sample index out of range";
      return PTYK_EINVAL;
    }
  hipSetDevice(h->device);
  double tot = 0;
  for (int64_t off = 0; off < count; off += h->n.batch) {
    int r = net_load_batch(h, order, off);
    if (r) return r;
    NCHK(h, nn_forward(h->n));
    int np = 0;
    NCHK(h, nn_loss(h->n, true, &np));
    NCHK(h, nn_backward(h->n));
    NCHK(h, nn_adam(h->n, lr));
    double l = 0;
    r = net_loss_value(h, np, &l);
    if (r) return r;
    tot += l;
  }
  if (mean_loss) *mean_loss = tot / (count / h->n.batch);
  return PTYK_OK;
}

int ptyk_net_evaluate(ptyk_net* h, const int64_t* order, int64_t count, double* mean_loss) {
  if (!h || !order || count <= 0 || count % h->n.batch || !mean_loss) return PTYK_EINVAL;
  for (int64_t i = 0; i < count; ++i)
    if (order[i] < 0 || order[i] >= h->ndata) {
      h->err = "sample index out of range";
      return PTYK_EINVAL;
    }
  hipSetDevice(h->device);
  double tot = 0;
  for (int64_t off = 0; off < count; off += h->n.batch) {
    int r = net_load_batch(h, order, off);
    if (r) return r;
    NCHK(h, nn_forward(h->n));
    int np = 0;
    NCHK(h, nn_loss(h->n, false, &np));
    double l = 0;
    r = net_loss_value(h, np, &l);
    if (r) return r;
    tot += l;
  }
  *mean_loss = tot / (count / h->n.batch);
  return PTYK_OK;
}

// Forward pass on host inputs (n x s x s float) -> outputs (n x 2 x s x s float); raw: the
// inputs are diffraction amplitudes, preprocessed on the device.
static int net_predict(ptyk_net* h, const float* inputs, int64_t n, float* outputs, bool raw) {
  if (!h || !inputs || !outputs || n <= 0) return PTYK_EINVAL;
  hipSetDevice(h->device);
  const long ss = (long)h->n.s * h->n.s;
  const int B = h->n.batch;
  float* stage = nullptr;
  NCHK(h, hipMalloc((void**)&stage, (size_t)B * ss * sizeof(float)));
  hipError_t e = hipSuccess;
  for (int64_t off = 0; off < n && e == hipSuccess; off += B) {
    const int64_t cnt = std::min<int64_t>(B, n - off);
    e = hipMemsetAsync(stage, 0, (size_t)B * ss * sizeof(float), h->n.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(stage, inputs + off * ss, (size_t)cnt * ss * sizeof(float), hipMemcpyHostToDevice,
                         h->n.stream);
    if (e == hipSuccess) {
      if (raw)
        hipLaunchKernelGGL(k_prep_amplitudes, dim3(B), dim3(256), 0, h->n.stream, stage, ss, h->n.bufs[B_X].a);
      else
        hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((B * ss + 255) / 256)), dim3(256), 0, h->n.stream, stage,
                           h->n.bufs[B_X].a, (long)(B * ss));
      e = nn_forward(h->n);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(outputs + off * 2 * ss, h->n.out, (size_t)cnt * 2 * ss * sizeof(float),
                         hipMemcpyDeviceToHost, h->n.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->n.stream);
  }
  hipFree(stage);
  NCHK(h, e);
  return PTYK_OK;
}

int ptyk_net_predict(ptyk_net* h, const float* inputs, int64_t n, float* outputs) {
  return net_predict(h, inputs, n, outputs, false);
}

int ptyk_net_predict_amplitudes(ptyk_net* h, const float* amplitudes, int64_t n, float* outputs) {
  return net_predict(h, amplitudes, n, outputs, true);
}

// Loss and parameter gradient of one batch (host inputs batch x s x s, targets batch x 2 x s x s);
// for gradient checks.
int ptyk_net_gradient(ptyk_net* h, const float* inputs, const float* targets, double* loss, float* grad) {
  if (!h || !inputs || !targets || !loss || !grad) return PTYK_EINVAL;
  hipSetDevice(h->device);
  const long ss = (long)h->n.s * h->n.s;
  const int B = h->n.batch;
  float* stage = nullptr;
  NCHK(h, hipMalloc((void**)&stage, (size_t)B * 2 * ss * sizeof(float)));
  hipError_t e = hipMemcpyAsync(stage, inputs, (size_t)B * ss * sizeof(float), hipMemcpyHostToDevice, h->n.stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((B * ss + 255) / 256)), dim3(256), 0, h->n.stream, stage,
                       h->n.bufs[B_X].a, (long)(B * ss));
    e = hipStreamSynchronize(h->n.stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(stage, targets, (size_t)B * 2 * ss * sizeof(float), hipMemcpyHostToDevice, h->n.stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((B * 2 * ss + 255) / 256)), dim3(256), 0, h->n.stream, stage,
                       h->n.target, (long)(B * 2 * ss));
    e = nn_forward(h->n);
  }
  int np = 0;
  if (e == hipSuccess) e = nn_loss(h->n, true, &np);
  if (e == hipSuccess) e = nn_backward(h->n);
  if (e == hipSuccess) e = hipStreamSynchronize(h->n.stream);
  hipFree(stage);
  NCHK(h, e);
  int r = net_loss_value(h, np, loss);
  if (r) return r;
  NCHK(h, hipMemcpyAsync(grad, h->n.g, h->n.nparam * sizeof(float), hipMemcpyDeviceToHost, h->n.stream));
  NCHK(h, hipStreamSynchronize(h->n.stream));
  return PTYK_OK;
}

}  // extern "C"
