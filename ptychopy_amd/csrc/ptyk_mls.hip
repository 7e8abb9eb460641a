// MLs: maximum-likelihood reconstruction with least-squares step sizes, gfx950 form.
//
// Follows MLS::iteration (MLS.cpp:74-490) group by group.  Every step of an iteration runs on
// the device -- including the per-frame 2x2 LSQ solves, the group means of the steps, the
// position refinement and the probe-evolution updates that the reference finishes on the host
// after ~130 synchronising reductions per group -- so an iteration is one stream of launches
// with no host round trip.  Scalars that a later launch needs live in MlsDev::gscal.
//
// Deviations from the reference's arithmetic, each deliberate:
//  * scatters (set_projections, d_addToArray_c with float atomics, utilitiesKernels.cu:1021-1062)
//    are gathers by canvas tile summing the group's frames in group order: deterministic;
//  * per-frame reductions accumulate in FP64 in a fixed order (thrust float sums in the
//    reference); the LSQ 2x2 systems are solved in FP64 (Eigen complex float there);
//  * the sub-pixel shift of the illumination probe e0*P0 + e1*V is applied as
//    IFFT(ramp * (e0 FFT(P0) + e1 FFT(V))) -- one inverse transform per frame instead of a
//    forward and an inverse (Probe::shift_probe on the combined probe, Probe.cpp:1261-1337);
//    equal up to FP32 rounding by linearity;
//  * 1/M^2 normalisations (powers of two) are folded into the spectral multiplies.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

namespace {

constexpr int NT = 256;
constexpr float kTwoPi = 6.283185307179586f;
// float(DBL_EPSILON): the imaginary offset MLS.cpp:301-304 adds to the object update
constexpr float kEpsF = 2.220446049250313e-16f;

__host__ __device__ inline int ppb_of(int m) { return m * m < 1024 ? m * m : 1024; }

// fftshifted frequency grid i/M - 0.5 rolled by M/2 (Probe.cpp:92-123 with d_shiftY)
__device__ __forceinline__ float freq(int i, int m) { return (float)((i + (m >> 1)) & (m - 1)) / (float)m - 0.5f; }

__device__ __forceinline__ float2 c_of(double re, double im) { return make_float2((float)re, (float)im); }

// Per-frame totals of chunk partials: one wave per frame, lanes over chunks, fixed order.
// tot[j * NV + v] = sum_c part[(j * ch + c) * NV + v].  A wave takes its two frames j and
// j + BT/64 together, every load issued before the first sum.  Ends with a block barrier.
template <int BT, int NV>
__device__ void frame_totals(const double* __restrict__ part, int ch, int G, double* __restrict__ tot) {
  constexpr int W = BT / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j0 = w; j0 < G; j0 += 2 * W) {
    const int j1 = j0 + W;
    const bool two = j1 < G;  // wave-uniform
    double s0[NV], s1[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s0[v] = s1[v] = 0.0;
    for (int c = lane; c < ch; c += 64) {
      const double* p0 = part + ((long)j0 * ch + c) * NV;
      const double* p1 = part + ((long)j1 * ch + c) * NV;
      double x0[NV], x1[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) x0[v] = p0[v];
#pragma unroll
      for (int v = 0; v < NV; ++v) x1[v] = two ? p1[v] : 0.0;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        s0[v] += x0[v];
        s1[v] += x1[v];
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) s0[v] = wave_sum(s0[v]);
    if (two) {
#pragma unroll
      for (int v = 0; v < NV; ++v) s1[v] = wave_sum(s1[v]);
    }
    if (lane == 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) tot[(long)j0 * NV + v] = s0[v];
      if (two) {
#pragma unroll
        for (int v = 0; v < NV; ++v) tot[(long)j1 * NV + v] = s1[v];
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------------ ROI origins
// IPtychoScanMesh::find_reconstruction_ROI (IPtychoScanMesh.cpp:123-172), in its float
// arithmetic: t = round(pos + ceil(Np_o/2 - M/2)), origin t - 1, sub = (col frac, row frac).
// An origin outside the canvas is clamped and flagged (the reference reads out of bounds).
__global__ void k_roi(const float2* __restrict__ pos, int n, int m, int rows, int cols, float cx, float cy,
                      int2* __restrict__ roi, float2* __restrict__ sub, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float p1 = pos[i].x + cx, p2 = pos[i].y + cy;
  const float t1 = roundf(p1), t2 = roundf(p2);
  sub[i] = make_float2(p2 - t2, p1 - t1);
  int ox = (int)t1 - 1, oy = (int)t2 - 1;
  if (!(t1 >= 1.f && t2 >= 1.f) || ox + m > rows || oy + m > cols) {
    flag[0] = 1;
    ox = min(max(ox, 0), rows - m);
    oy = min(max(oy, 0), cols - m);
  }
  roi[i] = make_int2(ox, oy);
}

// ------------------------------------------------------------------------ generic helpers
__global__ void k_scale_c(float2* __restrict__ x, long count, const double* __restrict__ s, int invert) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float f = invert ? (float)(1.0 / *s) : (float)(*s);
  x[i] = cscale(x[i], f);
}

__global__ void k_fill_c(float2* __restrict__ x, long count, float2 v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] = v;
}

// Canvas reductions into per-block partials: 0 = sum a, 1 = sum a |o|^2, 2 = max a,
// 3 = sum a^2 over a box, 4 = sum a^2 |o|^2 over a box (calc_object_norm).
__global__ __launch_bounds__(NT) void k_reduce(const float* __restrict__ a, const float2* __restrict__ o, long count,
                                               int op, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  __shared__ float redf[NT / 64];
  double acc = 0.0;
  float mx = 0.f;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < count; i += (long)gridDim.x * NT) {
    if (op == 0) acc += a[i];
    else if (op == 1) acc += (double)a[i] * ((double)o[i].x * o[i].x + (double)o[i].y * o[i].y);
    else mx = fmaxf(mx, a[i]);
  }
  if (op == 2) {
    mx = block_max<NT>(mx, redf);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
  } else {
    acc = block_sum<NT>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
  }
}

__global__ __launch_bounds__(NT) void k_box_partial(const float* __restrict__ w, const float2* __restrict__ o,
                                                    int cols, int x0, int y0, int by, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const long row = (long)(x0 + blockIdx.x) * cols + y0;
  double sw = 0.0, so = 0.0;
  for (int c = threadIdx.x; c < by; c += NT) {
    const double wv = w[row + c];
    const float2 ov = o[row + c];
    sw += wv * wv;
    so += ((double)ov.x * ov.x + (double)ov.y * ov.y) * wv * wv;
  }
  sw = block_sum<NT>(sw, red);
  so = block_sum<NT>(so, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sw;
    part[2 * blockIdx.x + 1] = so;
  }
}

// Probe::calc_object_norm (Probe.cpp:671-699): object_norm = rms(O W) / rms(W) over the box.
__global__ __launch_bounds__(NT) void k_objnorm(const double* __restrict__ part, int nb, double* __restrict__ gscal) {
  __shared__ double red[NT / 64];
  double sw = 0.0, so = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) {
    sw += part[2 * i];
    so += part[2 * i + 1];
  }
  sw = block_sum<NT>(sw, red);
  so = block_sum<NT>(so, red);
  if (threadIdx.x == 0) gscal[MLS_S_ONORM] = sqrt(so) / sqrt(sw);  // the box size cancels
}

__global__ __launch_bounds__(NT) void k_finish_max(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ float red[NT / 64];
  float mx = 0.f;
  for (int i = threadIdx.x; i < nb; i += NT) mx = fmaxf(mx, (float)part[i]);
  mx = block_max<NT>(mx, red);
  if (threadIdx.x == 0) *out = mx;
}

// IPtychoScanMesh::updateScanPositionsDiff (IPtychoScanMesh.cpp:340-369): float sums in
// position order (one thread, staged through LDS), then pos -= mean(pos) - mean(pos_o).
__global__ __launch_bounds__(NT) void k_pos_diff(float2* __restrict__ pos, const float2* __restrict__ pos_o, int n) {
  __shared__ float2 a[1024], b[1024];
  __shared__ float dxy[2];
  float s1x = 0.f, s1y = 0.f, s2x = 0.f, s2y = 0.f;
  for (int base = 0; base < n; base += 1024) {
    const int cnt = min(1024, n - base);
    for (int i = threadIdx.x; i < cnt; i += NT) {
      a[i] = pos[base + i];
      b[i] = pos_o[base + i];
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < cnt; ++i) {
        s1x += a[i].x;
        s1y += a[i].y;
        s2x += b[i].x;
        s2y += b[i].y;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float fn = (float)n;
    dxy[0] = s1x / fn - s2x / fn;
    dxy[1] = s1y / fn - s2y / fn;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NT) pos[i] = make_float2(pos[i].x - dxy[0], pos[i].y - dxy[1]);
}

// Probe::remove_extra_degree (Probe.cpp:615-669).  gscal[TMP] holds sum |V|^2.  The mean and
// the renormalisation of the evolution run over its first two entries (len = number of
// evolution rows in the reference).
__global__ __launch_bounds__(NT) void k_remove_extra(float2* __restrict__ vmode, float2* __restrict__ p0, int mm,
                                                     double* __restrict__ evo, int n,
                                                     const double* __restrict__ gscal) {
  __shared__ double cn_s;
  const float vn = sqrtf((float)(gscal[MLS_S_TMP] / mm));  // h_norm2: float result
  const float inv = (float)(1.0 / (double)vn);
  for (int q = threadIdx.x; q < mm; q += NT) vmode[q] = cscale(vmode[q], inv);
  for (int i = threadIdx.x; i < n; i += NT) evo[n + i] *= vn;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int len = n < 2 ? n : 2;
    double a = 0, b = 0;
    for (int i = 0; i < len; ++i) {
      a += evo[i];
      b += evo[n + i];
    }
    const double ma = (float)(a / len), mb = (float)(b / len);
    double s0 = 0;
    for (int i = 0; i < len; ++i) {
      evo[i] = (evo[i] - ma) * 0.99 + 1;
      evo[n + i] = evo[n + i] - mb;
      s0 += evo[i] * evo[i];
    }
    const double cn = sqrt(s0 / len);
    for (int i = 0; i < len; ++i) evo[i] /= cn;
    cn_s = cn;
  }
  __syncthreads();
  const float f = (float)cn_s;
  for (int q = threadIdx.x; q < mm; q += NT) p0[q] = cscale(p0[q], f);
}

// Illumination canvas (h_preCalillum / updated_cached_illumination, ProbeKernels.cu:691-730,
// Probe.cpp:701-729): s = sum over frames covering the pixel of |P0|^2 at the local offset, in
// frame order; mode 0 stores s, mode 1 the running average (p + s) / 2.  Tiles of 4 x 64; each
// chunk of 256 frames is first compacted (in frame order) to those whose ROI meets the tile.
__global__ __launch_bounds__(NT) void k_illum(float* __restrict__ illum, const float2* __restrict__ p0,
                                              const int2* __restrict__ roi, int n, int m, int rows, int cols,
                                              int mode) {
  __shared__ int2 lr[NT];
  __shared__ int wcnt[NT / 64];
  const int r0 = blockIdx.y * 4, c0 = blockIdx.x * 64;
  const int r = r0 + (threadIdx.x >> 6), c = c0 + (threadIdx.x & 63);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float s = 0.f;
  for (int base = 0; base < n; base += NT) {
    const int f = base + threadIdx.x;
    int2 o = make_int2(0, 0);
    bool hit = false;
    if (f < n) {
      o = roi[f];
      hit = !(o.x > r0 + 3 || o.x + m <= r0 || o.y > c0 + 63 || o.y + m <= c0);
    }
    const unsigned long long bal = __ballot(hit);
    __syncthreads();
    if (lane == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < NT / 64; ++i) {
      if (i < w) off += wcnt[i];
      tot += wcnt[i];
    }
    if (hit) lr[off + __popcll(bal & ((1ull << lane) - 1ull))] = o;
    __syncthreads();
    for (int j = 0; j < tot; ++j) {
      const int2 q = lr[j];
      const int lrw = r - q.x, lc = c - q.y;
      if (lrw >= 0 && lrw < m && lc >= 0 && lc < m) {
        const float2 v = p0[lrw * m + lc];
        const float a = sqrtf(v.x * v.x + v.y * v.y);
        s += a * a;
      }
    }
  }
  if (r < rows && c < cols) {
    const long i = (long)r * cols + c;
    illum[i] = mode ? (illum[i] + s) / 2.f : s;
  }
}

// ------------------------------------------------------------------------ one group
// Illumination probe (Probe::get_illumination_probe, Probe.cpp:887-954):
// vp_j = e0_j P0 + e1_j V, or with the sub-pixel shift its spectrum ramp_j (e0 F0 + e1 FV) / M^2
// (inverse-transformed by the caller).
__global__ void k_vp(float2* __restrict__ vp, const float2* __restrict__ spec, const float2* __restrict__ p0,
                     const float2* __restrict__ vmode, const double* __restrict__ evo, const float2* __restrict__ sub,
                     const int* __restrict__ gf, int n, int m, int subpix, int var_any) {
  const int mm = m * m;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const int j = blockIdx.y, f = gf[j];
  const float e0 = var_any ? (float)evo[f] : 1.f, e1 = var_any ? (float)evo[n + f] : 0.f;
  float2 v;
  if (subpix) {
    const float2 a = spec[q], b = spec[mm + q];
    v = make_float2(e0 * a.x + e1 * b.x, e0 * a.y + e1 * b.y);
    const int r = q / m, c = q % m;
    const float2 s = sub[f];
    const float nom = s.x * freq(c, m) + s.y * freq(r, m);
    float sn, cs;
    sincosf(nom * -kTwoPi, &sn, &cs);
    v = cscale(cmul(v, make_float2(cs, sn)), 1.f / (float)mm);
  } else {
    const float2 a = p0[q], b = vmode[q];
    v = make_float2(e0 * a.x + e1 * b.x, e0 * a.y + e1 * b.y);
  }
  vp[(long)j * mm + q] = v;
}

// The sub-pixel form of k_vp in three launches, column pass first: rspec holds the row spectra
// R_P = FFT_rows(P_0) and R_V = FFT_rows(V) (fft_rows_pair); per frame a column pass transforms
// e0 R_P + e1 R_V along the columns, multiplies by the row-frequency ramp exp(-2 pi i sy f_row) / M
// and transforms back, then a row pass multiplies by the column-frequency ramp
// exp(-2 pi i sx f_col) / M and inverse-transforms the rows.  k_vp's 2D ramp is the product of the
// two and transforms along rows and columns commute, so this is
// IFFT2(ramp_j (e0 FFT2(P_0) + e1 FFT2(V))) / M^2 (Probe::shift_probe of the frame's probe), in
// another rounding order than 2D spectra, a ramped inverse row pass and an inverse column pass (one
// launch more, which measured 0.5 ms per 51x51 / 256^2 / 5-mode iteration slower).
template <int M>
__global__ __launch_bounds__(NT) void k_vp_cols(float2* __restrict__ vp, const float2* __restrict__ rspec,
                                                const double* __restrict__ evo, const float2* __restrict__ sub,
                                                const int* __restrict__ gf, int n, int var_any,
                                                const float2* __restrict__ tw) {
  constexpr int NC = (plan_min_batch<M, NT>() + 15) / 16 * 16;
  constexpr int PAIRS = NC * M / 2 / NT;
  constexpr int MM = M * M;
  static_assert((NC * M / 2) % NT == 0, "column tile must divide among the threads");
  __shared__ float2 lds[NC * M];
  const int tid = threadIdx.x;
  const int j = blockIdx.y, f = gf[j];
  const float e0 = var_any ? (float)evo[f] : 1.f, e1 = var_any ? (float)evo[n + f] : 0.f;
  const float sy = sub[f].y;
  auto off = [&](int e) {  // e: strip-major element index of the tile (even)
    return (long)((e >> 4) % M) * M + (e / (M * 16)) * 16 + (e & 15);
  };
  const float2* ra = rspec + blockIdx.x * NC;
  {
    float4 av[PAIRS], bv[PAIRS];
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) {
      const long q = off(2 * (tid + k * NT));
      av[k] = *reinterpret_cast<const float4*>(ra + q);
      bv[k] = *reinterpret_cast<const float4*>(ra + MM + q);
    }
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) {
      const float4 a = av[k], b = bv[k];
      *reinterpret_cast<float4*>(&lds[2 * (tid + k * NT)]) =
          make_float4(e0 * a.x + e1 * b.x, e0 * a.y + e1 * b.y, e0 * a.z + e1 * b.z, e0 * a.w + e1 * b.w);
    }
  }
  __syncthreads();
  fft_lds<M, false, NC, NT, StripLayout<M>, true>(lds, tw, tid);
  const float inv_m = 1.f / (float)M;
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const int e = 2 * (tid + k * NT);
    float sn, cs;
    sincosf(sy * freq((e >> 4) % M, M) * -kTwoPi, &sn, &cs);  // both values of the pair share the row
    float4* p = reinterpret_cast<float4*>(&lds[e]);
    const float4 t = *p;
    const float2 u0 = cscale(cmul(make_float2(t.x, t.y), make_float2(cs, sn)), inv_m);
    const float2 u1 = cscale(cmul(make_float2(t.z, t.w), make_float2(cs, sn)), inv_m);
    *p = make_float4(u0.x, u0.y, u1.x, u1.y);
  }
  __syncthreads();
  fft_lds<M, true, NC, NT, StripLayout<M>, true>(lds, tw, tid);
  float2* out = vp + (long)j * MM + blockIdx.x * NC;
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const int e = 2 * (tid + k * NT);
    *reinterpret_cast<float4*>(out + off(e)) = *reinterpret_cast<const float4*>(&lds[e]);
  }
}

template <int M>
__global__ __launch_bounds__(NT) void k_vp_rows_ramp(float2* __restrict__ vp, const float2* __restrict__ sub,
                                                     const int* __restrict__ gf, const float2* __restrict__ tw) {
  constexpr int NB = plan_min_batch<M, NT>();
  constexpr int PAIRS = NB * M / 2 / NT;
  static_assert(M % NB == 0 && (NB * M / 2) % NT == 0, "a workgroup's rows must share a plane");
  static_assert((2 * NT) % M == 0, "a thread's columns must repeat across its rows");
  using L = SwzRowLayout<M, M>;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * NB;
  const float sx = sub[gf[(int)(row0 / M)]].x;
  const int c = (2 * tid) % M;  // the thread's two columns, the same in every row it holds
  const float inv_m = 1.f / (float)M;
  float2 w0, w1;
  sincosf(sx * freq(c, M) * -kTwoPi, &w0.y, &w0.x);
  sincosf(sx * freq(c + 1, M) * -kTwoPi, &w1.y, &w1.x);
  {
    float4 v[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) v[i] = *reinterpret_cast<const float4*>(vp + row0 * M + 2 * (tid + i * NT));
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (tid + i * NT);
      const float2 u0 = cscale(cmul(make_float2(v[i].x, v[i].y), w0), inv_m);
      const float2 u1 = cscale(cmul(make_float2(v[i].z, v[i].w), w1), inv_m);
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) = make_float4(u0.x, u0.y, u1.x, u1.y);
    }
  }
  __syncthreads();
  fft_lds<M, true, NB, NT, L, false>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * NT);
    *reinterpret_cast<float4*>(vp + row0 * M + e) = *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
  }
}

// The illumination probes of G frames from the row spectra rspec (see k_vp_cols).
static hipError_t vp_from_row_spectra(float2* vp, const float2* rspec, const double* evo, const float2* sub,
                                      const int* gf, int n, int var_any, int G, int m, const float2* tw,
                                      hipStream_t s) {
  switch (m) {
#define PTYK_VPC_CASE(MM)                                                                                         \
  case MM: {                                                                                                      \
    constexpr int NC = (plan_min_batch<MM, NT>() + 15) / 16 * 16;                                                \
    hipLaunchKernelGGL((k_vp_cols<MM>), dim3(MM / NC, G), dim3(NT), 0, s, vp, rspec, evo, sub, gf, n, var_any,   \
                       tw);                                                                                       \
    hipLaunchKernelGGL((k_vp_rows_ramp<MM>), dim3((unsigned)((long)G * MM / plan_min_batch<MM, NT>())), dim3(NT), \
                       0, s, vp, sub, gf, tw);                                                                    \
    return hipGetLastError();                                                                                     \
  }
    PTYK_VPC_CASE(128)
    PTYK_VPC_CASE(256)
    PTYK_VPC_CASE(512)
#undef PTYK_VPC_CASE
    default: return hipErrorInvalidValue;
  }
}

// obj_proj = O at the ROI (d_readFromArray_c, utilitiesKernels.cu:1064-1084);
// psi_0 = obj_proj * vp (with the sub-pixel shift) or P0 * obj_proj, psi_k = P_k * obj_proj.
__global__ void k_psi(float2* __restrict__ psi, float2* __restrict__ op, float2* __restrict__ op2,
                      const float2* __restrict__ vp, const float2* __restrict__ obj, const float2* __restrict__ probe,
                      const int2* __restrict__ roi, const int* __restrict__ gf, int G, int K, int m, int cols,
                      int subpix) {
  const int mm = m * m;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const int j = blockIdx.y;
  const int2 o = roi[gf[j]];
  const int r = q / m, c = q % m;
  const float2 O = obj[(long)(o.x + r) * cols + o.y + c];
  op[(long)j * mm + q] = O;
  if (op2) op2[(long)j * mm + q] = O;  // input of the position solver's transform
  psi[(long)j * mm + q] = subpix ? cmul(O, vp[(long)j * mm + q]) : cmul(probe[q], O);
  for (int k = 1; k < K; ++k) psi[((long)k * G + j) * mm + q] = cmul(probe[(long)k * mm + q], O);
}

// Fourier error partials and the modulus constraint (Diffractions::get_fourier_error and
// modulus_constraint, Diffractions.cpp:261-310; d_modulus_amplitude DiffractionsKernels.cu:127-137):
// R = A / (|Psi| + 1e-9) - 1 (the 1e-9 literal makes it a double expression), Psi_k *= R / M^2.
__global__ __launch_bounds__(NT) void k_modulus(float2* __restrict__ psi, const __half* __restrict__ amp,
                                                const float* __restrict__ amp_scale, const int* __restrict__ gf,
                                                int G, int K, int m, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int mm = m * m, ppb = ppb_of(m);
  const int j = blockIdx.y, f = gf[j];
  const float sc = amp_scale[f];
  const float inv_mm = 1.f / (float)mm;
  double e = 0.0;
  for (int t = threadIdx.x; t < ppb; t += NT) {
    const int q = blockIdx.x * ppb + t;
    float s2 = 0.f;
    for (int k = 0; k < K; ++k) s2 += cabs2(psi[((long)k * G + j) * mm + q]);
    const float ap = sqrtf(s2);
    const float A = load_amp(amp, (long)f * mm + q, sc);
    const double d = (double)(A - ap);
    e += d * d * 4.0;
    const float R = (float)((double)A / ((double)ap + 1e-9) - 1.0) * inv_mm;
    for (int k = 0; k < K; ++k) {
      const long i = ((long)k * G + j) * mm + q;
      psi[i] = cscale(psi[i], R);
    }
  }
  e = block_sum<NT>(e, red);
  if (threadIdx.x == 0) part[(long)j * gridDim.x + blockIdx.x] = e;
}

// err_j = sqrt(mean((A - |Psi|)^2 / 0.25)) (the division by 0.25 doubles the RMS).
__global__ void k_ferr(const double* __restrict__ part, int ch, const int* __restrict__ gf, int G, int mm,
                       float* __restrict__ ferr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= G) return;
  double s = 0.0;
  for (int c = 0; c < ch; ++c) s += part[(long)j * ch + c];
  ferr[gf[j]] = (float)sqrt(s / mm);
}

// probe_update_j = chi_0 conj(obj_proj_j) (h_multiplyConju).
__global__ void k_pu(float2* __restrict__ pu, const float2* __restrict__ chi, const float2* __restrict__ op, long count) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const float2 a = chi[i], b = op[i];
    pu[i] = make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
  }
}

// x_j *= exp(-2 pi i sign (sub.x f[col] + sub.y f[row])) / M^2 (spectral shift, 1/M^2 folded).
__global__ void k_ramp(float2* __restrict__ x, const float2* __restrict__ sub, const int* __restrict__ gf, int m,
                       float sign) {
  const int mm = m * m;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const int j = blockIdx.y;
  const float2 s = sub[gf[j]];
  const int r = q / m, c = q % m;
  const float nom = (sign * s.x) * freq(c, m) + (sign * s.y) * freq(r, m);
  float sn, cs;
  sincosf(nom * -kTwoPi, &sn, &cs);
  const long i = (long)j * mm + q;
  x[i] = cscale(cmul(x[i], make_float2(cs, sn)), 1.f / (float)mm);
}

// probe_update_m = mean_j (p / (p + 0.01 MAX_ILLUM))_roi_j probe_update_j  (MLS.cpp:251-279,
// d_addFactorDivide utilitiesKernels.cu:171-179).  on_the_fly: probe_update_j = chi conj(op_j).
__device__ __forceinline__ void pum_block(int block, float2* __restrict__ pum, const float2* __restrict__ src,
                                          const float2* __restrict__ op, const float* __restrict__ illum,
                                          const int2* __restrict__ roi, const int* __restrict__ gf, int G, int m,
                                          int cols, float fw, int on_the_fly) {
  const int mm = m * m;
  const int q = block * NT + threadIdx.x;
  if (q >= mm) return;
  const int r = q / m, c = q % m;
  float2 acc = make_float2(0.f, 0.f);
  auto add = [&](float p, float2 u) {
    const float w = p / (p + fw);
    acc.x += w * u.x;
    acc.y += w * u.y;
  };
  int j = 0;
  if (!on_the_fly) {
    // eight frames' loads in flight before their sums (which stay in frame order)
    constexpr int B = 8;
    for (; j + B <= G; j += B) {
      float p[B];
      float2 u[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int2 o = roi[gf[j + b]];
        p[b] = illum[(long)(o.x + r) * cols + o.y + c];
        u[b] = src[(long)(j + b) * mm + q];
      }
#pragma unroll
      for (int b = 0; b < B; ++b) add(p[b], u[b]);
    }
  }
#pragma unroll 4
  for (; j < G; ++j) {
    const int2 o = roi[gf[j]];
    const float p = illum[(long)(o.x + r) * cols + o.y + c];
    float2 u = src[(long)j * mm + q];
    if (on_the_fly) {
      const float2 b = op[(long)j * mm + q];
      u = make_float2(u.x * b.x + u.y * b.y, u.y * b.x - u.x * b.y);
    }
    add(p, u);
  }
  pum[q] = cscale(acc, (float)(1.0 / G));
}

__global__ void k_pum(float2* __restrict__ pum, const float2* __restrict__ src,
                                            const float2* __restrict__ op, const float* __restrict__ illum,
                                            const int2* __restrict__ roi, const int* __restrict__ gf, int G, int m,
                                            int cols, float fw, int on_the_fly) {
  pum_block(blockIdx.x, pum, src, op, illum, roi, gf, G, m, cols, fw, on_the_fly);
}

// Probe updates of modes 2..K in one launch (MLS.cpp:251-279 for ll >= 2, no sub-pixel shift):
// pum_ll = mean_j w_j chi_ll,j conj(op_j), P_ll += pum_ll * mean probe step, one thread per
// (pixel, mode) with the frames summed in group order.  Nothing else in the group reads these
// modes between the reference's per-mode passes, so they fuse exactly.
__global__ void k_pum_modes(float2* __restrict__ probe, const float2* __restrict__ psi, const float2* __restrict__ op,
                            const float* __restrict__ illum, const int2* __restrict__ roi, const int* __restrict__ gf,
                            int G, int m, int cols, float fw, const double* __restrict__ gscal) {
  const int mm = m * m;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const int k = blockIdx.y + 1;
  const int r = q / m, c = q % m;
  float2 acc = make_float2(0.f, 0.f);
  const float2* chi = psi + (long)k * G * mm;
  for (int j = 0; j < G; ++j) {
    const int2 o = roi[gf[j]];
    const float p = illum[(long)(o.x + r) * cols + o.y + c];
    const float w = p / (p + fw);
    const float2 b = op[(long)j * mm + q];
    const float2 u = chi[(long)j * mm + q];
    const float2 d = make_float2(u.x * b.x + u.y * b.y, u.y * b.x - u.x * b.y);
    acc.x += w * d.x;
    acc.y += w * d.y;
  }
  const float inv = (float)(1.0 / G);
  const float2 bp = c_of(gscal[MLS_S_BP], gscal[MLS_S_BP + 1]);
  const long i = (long)k * mm + q;
  probe[i] = cadd(probe[i], cmul(cscale(acc, inv), bp));
}

// Object update direction of frame j at local pixel q: chi conj(vp_j) for the first probe mode,
// P_llp chi (no conjugate, MLS.cpp:294-297) for the others.
__device__ __forceinline__ float2 obj_dir(const float2* __restrict__ chi, const float2* __restrict__ vp,
                                          const float2* __restrict__ pl, long i, int q) {
  const float2 a = chi[i];
  if (pl == nullptr) {
    const float2 b = vp[i];
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
  }
  return cmul(pl[q], a);
}

// object_upd_sum = (0, eps) + set_projections(direction), preconditioned by
// 1 / (p + delta_p MAX_ILLUM) (d_object_sum_update_Gfun, utilitiesKernels.cu:181-189).
// Gather by canvas tile of (4 RPT) x 64 pixels, RPT rows per thread, over the group's frames
// in group order.
constexpr int PRE_RPT = 2;
template <int RPT>
__device__ __forceinline__ void precond_tile(int bx, int by, float2* __restrict__ pre, const float* __restrict__ illum,
                                             const float2* __restrict__ chi, const float2* __restrict__ vp,
                                             const float2* __restrict__ pl, const int2* __restrict__ roi,
                                             const int* __restrict__ gf, int G, int m, int rows, int cols, float fd,
                                             int scatter) {
  constexpr int TH = 4 * RPT;
  __shared__ int2 lr[NT];
  const int mm = m * m;
  const int r0 = by * TH, c0 = bx * 64;
  const int ty = threadIdx.x >> 6, c = c0 + (threadIdx.x & 63);
  float2 u[RPT];
  float il[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    u[k] = make_float2(0.f, kEpsF);
    const int r = r0 + ty + 4 * k;
    il[k] = scatter && r < rows && c < cols ? illum[(long)r * cols + c] : 0.f;
  }
  if (scatter) {
    for (int base = 0; base < G; base += NT) {
      const int cnt = min(NT, G - base);
      __syncthreads();
      if ((int)threadIdx.x < cnt) lr[threadIdx.x] = roi[gf[base + threadIdx.x]];
      __syncthreads();
      // frames whose ROI meets the tile (a block-uniform test), four at a time: their loads are
      // issued before the in-order accumulation
      constexpr int U = 4;
      int pend[U];
      int np = 0;
      auto flush = [&]() {
        float2 d[U][RPT];
        bool in[U][RPT];
#pragma unroll
        for (int t = 0; t < U; ++t) {
          const int2 o = lr[pend[t < np ? t : 0]];
          const int lc = c - o.y;
          const bool col_in = t < np && lc >= 0 && lc < m;
#pragma unroll
          for (int k = 0; k < RPT; ++k) {
            const int lrw = r0 + ty + 4 * k - o.x;
            in[t][k] = col_in && lrw >= 0 && lrw < m;
            if (in[t][k]) {
              const int q = lrw * m + lc;
              d[t][k] = obj_dir(chi, vp, pl, (long)(base + pend[t]) * mm + q, q);
            }
          }
        }
#pragma unroll
        for (int t = 0; t < U; ++t)
#pragma unroll
          for (int k = 0; k < RPT; ++k)
            if (in[t][k]) u[k] = cadd(u[k], d[t][k]);
        np = 0;
      };
      for (int jj = 0; jj < cnt; ++jj) {
        const int2 o = lr[jj];
        if (o.x > r0 + TH - 1 || o.x + m <= r0 || o.y > c0 + 63 || o.y + m <= c0) continue;  // block-uniform
        pend[np++] = jj;
        if (np == U) flush();
      }
      if (np > 0) flush();
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = r0 + ty + 4 * k;
    if (r < rows && c < cols) {
      float2 v = u[k];
      if (scatter) {
        const float dn = il[k] + fd;
        v = make_float2(v.x / dn, v.y / dn);
      }
      pre[(long)r * cols + c] = v;
    }
  }
}

template <int RPT>
__global__ __launch_bounds__(NT) void k_precond(float2* __restrict__ pre, const float* __restrict__ illum,
                                                const float2* __restrict__ chi, const float2* __restrict__ vp,
                                                const float2* __restrict__ pl, const int2* __restrict__ roi,
                                                const int* __restrict__ gf, int G, int m, int rows, int cols, float fd,
                                                int scatter) {
  precond_tile<RPT>(blockIdx.x, blockIdx.y, pre, illum, chi, vp, pl, roi, gf, G, m, rows, cols, fd, scatter);
}

// The probe-update average (mode 1, from the shifted probe updates) and the object direction in
// one launch: they read disjoint inputs and write disjoint outputs.  The first n_pum blocks
// (one frame loop per probe pixel, the long ones) are dispatched first; the canvas tiles, short,
// fill the rest of the chip behind them.  Each block's arithmetic is that of k_pum / k_precond.
template <int RPT>
__global__ __launch_bounds__(NT) void k_pum_precond(float2* __restrict__ pum, const float2* __restrict__ pu,
                                                    float fw, int n_pum, float2* __restrict__ pre,
                                                    const float* __restrict__ illum, const float2* __restrict__ chi,
                                                    const float2* __restrict__ vp, const float2* __restrict__ pl,
                                                    const int2* __restrict__ roi, const int* __restrict__ gf, int G,
                                                    int m, int rows, int cols, float fd, int scatter, int tiles_x) {
  const int b = blockIdx.x;
  if (b < n_pum) {
    pum_block(b, pum, pu, nullptr, illum, roi, gf, G, m, cols, fw, 0);
  } else {
    const int t = b - n_pum;
    precond_tile<RPT>(t % tiles_x, t / tiles_x, pre, illum, chi, vp, pl, roi, gf, G, m, rows, cols, fd, scatter);
  }
}

__global__ void k_dir(float2* __restrict__ out, const float2* __restrict__ chi, const float2* __restrict__ vp,
                      const float2* __restrict__ pl, int m) {
  const int mm = m * m;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const long i = (long)blockIdx.y * mm + q;
  out[i] = obj_dir(chi, vp, pl, i, q);
}

// LSQ sums per frame chunk (d_get_optimal_step_lsq, utilitiesKernels.cu:445-466), FP64:
// [|dOP|^2, Re dOP conj dPO, Im dOP conj dPO, |dPO|^2, Re conj(dOP) chi, Re conj(dPO) chi].
// The object direction dO is read from the canvas at the frame's ROI when canvas is given
// (get_projections of the preconditioned update), else from oup frame by frame.
__global__ __launch_bounds__(NT) void k_lsq_partial(const float2* __restrict__ oup, const float2* __restrict__ canvas,
                                                    const int2* __restrict__ roi, const int* __restrict__ gf,
                                                    int cols, const float2* __restrict__ vp,
                                                    const float2* __restrict__ pum, const float2* __restrict__ op,
                                                    const float2* __restrict__ chi, int m, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int mm = m * m, ppb = ppb_of(m);
  const int j = blockIdx.y;
  const int2 o = canvas ? roi[gf[j]] : make_int2(0, 0);
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < ppb; t += NT) {
    const int q = blockIdx.x * ppb + t;
    const long i = (long)j * mm + q;
    const float2 a = canvas ? canvas[(long)(o.x + q / m) * cols + o.y + q % m] : oup[i];
    const float2 b = vp[i], p = pum[q], o2 = op[i], x = chi[i];
    const double opr = (double)a.x * b.x - (double)a.y * b.y, opi = (double)a.x * b.y + (double)a.y * b.x;
    const double por = (double)p.x * o2.x - (double)p.y * o2.y, poi = (double)p.x * o2.y + (double)p.y * o2.x;
    s[0] += opr * opr + opi * opi;
    s[1] += opr * por + opi * poi;
    s[2] += opi * por - opr * poi;
    s[3] += por * por + poi * poi;
    s[4] += opr * x.x + opi * x.y;
    s[5] += por * x.x + poi * x.y;
  }
  for (int v = 0; v < 6; ++v) {
    const double t = block_sum<NT>(s[v], red);
    if (threadIdx.x == 0) part[((long)j * gridDim.x + blockIdx.x) * 6 + v] = t;
  }
}

// Per frame: AA with the 0.1 x group-mean diagonal load, 2x2 solve in FP64, steps scaled by
// beta * beta_LSQ (MLS.cpp:370-416); then the group means of the object and probe steps
// (Sample::update_object with one scan id, Sample.cpp:277-331; Probe::update_probe,
// Probe.cpp:1339-1364).  solve = 0: only the means of the current steps.
__global__ __launch_bounds__(1024) void k_lsq_solve(const double* __restrict__ part, int ch,
                                                    const int* __restrict__ gf, int G, double lam_mm, double c_obj,
                                                    double c_probe, int solve, float2* __restrict__ beta_obj,
                                                    float2* __restrict__ beta_probe, double* __restrict__ ftot,
                                                    double* __restrict__ gscal) {
  // after the frame totals one wave does the rest: a group has few frames, and wave
  // reductions need no block barrier
  if (solve) frame_totals<1024, 6>(part, ch, G, ftot);
  if (threadIdx.x >= 64) return;
  const int ln = threadIdx.x;
  double sm[4] = {0, 0, 0, 0};
  if (solve) {
    double a1 = 0.0, a4 = 0.0;
    for (int j = ln; j < G; j += 64) {
      a1 += ftot[j * 6 + 0] + lam_mm;
      a4 += ftot[j * 6 + 3] + lam_mm;
    }
    const double m1 = 0.1 * wave_sum(a1) / G;
    const double m4 = 0.1 * wave_sum(a4) / G;
    for (int j = ln; j < G; j += 64) {
      const double* t = ftot + j * 6;
      const double A11 = t[0] + lam_mm + m1, A22 = t[3] + lam_mm + m4;
      const double a2r = t[1], a2i = t[2], b1 = t[4], b2 = t[5];
      // A = [[A11, a2], [conj(a2), A22]], det real
      const double det = A11 * A22 - (a2r * a2r + a2i * a2i);
      const double x0r = (b1 * A22 - a2r * b2) / det, x0i = (-a2i * b2) / det;
      const double x1r = (A11 * b2 - a2r * b1) / det, x1i = (a2i * b1) / det;
      const float2 bo = c_of(c_obj * x0r, c_obj * x0i), bp = c_of(c_probe * x1r, c_probe * x1i);
      beta_obj[gf[j]] = bo;
      beta_probe[gf[j]] = bp;
      sm[0] += bo.x;  // the means take the stored (float) steps
      sm[1] += bo.y;
      sm[2] += bp.x;
      sm[3] += bp.y;
    }
  } else {
    for (int j = ln; j < G; j += 64) {
      const float2 bo = beta_obj[gf[j]], bp = beta_probe[gf[j]];
      sm[0] += bo.x;
      sm[1] += bo.y;
      sm[2] += bp.x;
      sm[3] += bp.y;
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double t = wave_sum(sm[v]);
    if (ln == 0) gscal[MLS_S_BO + v] = t / G;
  }
}

// x += y * beta, beta from gscal (object: O += precond * mean step; probe: P_ll += pum * mean step);
// the second update (x2, y2, count2, beta2), when count2 > 0, runs in the same launch after the
// first's range.  Each thread takes two consecutive elements of one range with 16-byte loads and
// stores (both bases are 16-byte aligned: device allocations and whole probe modes).
__device__ __forceinline__ void axpy_pair(float2* __restrict__ x, const float2* __restrict__ y, long e, long count,
                                          float2 b) {
  if (e + 1 < count) {
    const float4 xv = *reinterpret_cast<const float4*>(x + e);
    const float4 yv = *reinterpret_cast<const float4*>(y + e);
    const float2 r0 = cadd(make_float2(xv.x, xv.y), cmul(make_float2(yv.x, yv.y), b));
    const float2 r1 = cadd(make_float2(xv.z, xv.w), cmul(make_float2(yv.z, yv.w), b));
    *reinterpret_cast<float4*>(x + e) = make_float4(r0.x, r0.y, r1.x, r1.y);
  } else if (e < count) {
    x[e] = cadd(x[e], cmul(y[e], b));
  }
}

__global__ void k_axpy_beta(float2* __restrict__ x, const float2* __restrict__ y, long count,
                            const double* __restrict__ beta, float2* __restrict__ x2, const float2* __restrict__ y2,
                            long count2, const double* __restrict__ beta2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long h1 = (count + 1) / 2;
  if (t < h1)
    axpy_pair(x, y, 2 * t, count, c_of(beta[0], beta[1]));
  else
    axpy_pair(x2, y2, 2 * (t - h1), count2, c_of(beta2[0], beta2[1]));
}

// Gradient position solver spectra (Probe.cpp:956-1118): t = FFT(obj_proj) * 2 pi i;
// x <- t f[col] / M^2 (d/dcol), y <- t f[row] / M^2 (d/drow).
__global__ void k_deriv(float2* __restrict__ x, float2* __restrict__ y, int m) {
  const int mm = m * m;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const long i = (long)blockIdx.y * mm + q;
  const float2 f = x[i];
  const float2 t = make_float2(-f.y * kTwoPi, f.x * kTwoPi);
  const float inv = 1.f / (float)mm;
  x[i] = cscale(t, freq(q % m, m) * inv);
  y[i] = cscale(t, freq(q / m, m) * inv);
}

// [sum |dX vp|^2, sum Re(chi conj(dX vp)), same for dY] per frame chunk.
__global__ __launch_bounds__(NT) void k_pos_partial(const float2* __restrict__ dx, const float2* __restrict__ dy,
                                                    const float2* __restrict__ vp, const float2* __restrict__ chi, int m,
                                                    double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int mm = m * m, ppb = ppb_of(m);
  const int j = blockIdx.y;
  double s[4] = {0, 0, 0, 0};
  for (int t = threadIdx.x; t < ppb; t += NT) {
    const long i = (long)j * mm + blockIdx.x * ppb + t;
    const float2 v = vp[i], x = chi[i], a = dx[i], b = dy[i];
    const double axr = (double)a.x * v.x - (double)a.y * v.y, axi = (double)a.x * v.y + (double)a.y * v.x;
    const double ayr = (double)b.x * v.x - (double)b.y * v.y, ayi = (double)b.x * v.y + (double)b.y * v.x;
    s[0] += axr * axr + axi * axi;
    s[1] += x.x * axr + x.y * axi;
    s[2] += ayr * ayr + ayi * ayi;
    s[3] += x.x * ayr + x.y * ayi;
  }
  for (int v = 0; v < 4; ++v) {
    const double t = block_sum<NT>(s[v], red);
    if (threadIdx.x == 0) part[((long)j * gridDim.x + blockIdx.x) * 4 + v] = t;
  }
}

// Position step from one-dimensional transforms.  The gradient solver's derivative images
// dX = IFFT2(FFT2(O_r) 2 pi i f_col) / M^2 and dY = IFFT2(FFT2(O_r) 2 pi i f_row) / M^2
// (Probe.cpp:956-1118) have multipliers that depend on one frequency index only, so the
// transforms along the other axis cancel (IFFT o FFT = M along an axis):
// dX = IFFT_rows(FFT_rows(O_r) 2 pi i f_col) / M and dY = IFFT_cols(FFT_cols(O_r) 2 pi i f_row) / M.
// The exit-wave row pass (k_mls_pass_a) writes dX while the patch rows are in its LDS; k_pos_cols
// forms dY of whole 16-column strips in LDS and sums the four products with vp and chi over
// them, so dY is never stored.

// t 2 pi i s for a spectrum value t and a real scale s
__device__ __forceinline__ float2 deriv_of(float2 t, float sc) {
  return cscale(make_float2(-t.y * kTwoPi, t.x * kTwoPi), sc);
}

// Partials part[(j * gridDim.x + blockIdx.x) * 4 + v] of [sum |dX vp|^2, sum Re(chi conj(dX vp)),
// same for dY] over the workgroup's NC columns of frame j.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_pos_cols(const float2* __restrict__ op, const float2* __restrict__ dx,
                                                      const float2* __restrict__ vp, const float2* __restrict__ chi,
                                                      const float2* __restrict__ tw, double* __restrict__ part) {
  constexpr int NC = (plan_min_batch<M, PTYK_NT>() + 15) / 16 * 16;
  constexpr int PAIRS = NC * M / 2 / PTYK_NT;
  constexpr bool PRE = PAIRS <= 8;  // dX, vp and chi in flight during the transforms
  static_assert((NC * M / 2) % PTYK_NT == 0, "column tile must divide among the threads");
  __shared__ float2 lds[NC * M];
  const int tid = threadIdx.x;
  const int j = blockIdx.y, c0 = blockIdx.x * NC;
  const long fbase = (long)j * M * M + c0;
  auto gaddr = [&](int e) {  // e: strip-major element index of the tile (even)
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    return fbase + (long)i * M + g * 16 + c;
  };
  {
    float4 o[PAIRS];
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) o[k] = *reinterpret_cast<const float4*>(op + gaddr(2 * (tid + k * PTYK_NT)));
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) *reinterpret_cast<float4*>(&lds[2 * (tid + k * PTYK_NT)]) = o[k];
  }
  float4 a[PAIRS], v[PAIRS], x[PAIRS];
  auto load_rest = [&]() {
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) {
      const long q = gaddr(2 * (tid + k * PTYK_NT));
      a[k] = *reinterpret_cast<const float4*>(dx + q);
      v[k] = *reinterpret_cast<const float4*>(vp + q);
      x[k] = *reinterpret_cast<const float4*>(chi + q);
    }
  };
  if constexpr (PRE) load_rest();
  __syncthreads();
  fft_lds<M, false, NC, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
  const float inv = 1.f / (float)M;
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const int e = 2 * (tid + k * PTYK_NT);
    const float sc = freq((e >> 4) % M, M) * inv;  // both values of the pair share row (e >> 4) % M
    float4* p = reinterpret_cast<float4*>(&lds[e]);
    const float4 t = *p;
    const float2 b0 = deriv_of(make_float2(t.x, t.y), sc), b1 = deriv_of(make_float2(t.z, t.w), sc);
    *p = make_float4(b0.x, b0.y, b1.x, b1.y);
  }
  __syncthreads();
  fft_lds<M, true, NC, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
  if constexpr (!PRE) load_rest();
  double s[4] = {0, 0, 0, 0};
  auto acc = [&](float2 ax, float2 ay, float2 vv, float2 xx) {
    const double axr = (double)ax.x * vv.x - (double)ax.y * vv.y, axi = (double)ax.x * vv.y + (double)ax.y * vv.x;
    const double ayr = (double)ay.x * vv.x - (double)ay.y * vv.y, ayi = (double)ay.x * vv.y + (double)ay.y * vv.x;
    s[0] += axr * axr + axi * axi;
    s[1] += xx.x * axr + xx.y * axi;
    s[2] += ayr * ayr + ayi * ayi;
    s[3] += xx.x * ayr + xx.y * ayi;
  };
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const float4 b = *reinterpret_cast<const float4*>(&lds[2 * (tid + k * PTYK_NT)]);
    acc(make_float2(a[k].x, a[k].y), make_float2(b.x, b.y), make_float2(v[k].x, v[k].y),
        make_float2(x[k].x, x[k].y));
    acc(make_float2(a[k].z, a[k].w), make_float2(b.z, b.w), make_float2(v[k].z, v[k].w),
        make_float2(x[k].z, x[k].w));
  }
  double* red = reinterpret_cast<double*>(lds);  // block_sum syncs before its first write
  for (int t = 0; t < 4; ++t) {
    const double r = block_sum<PTYK_NT>(s[t], red);
    if (tid == 0) part[((long)j * gridDim.x + blockIdx.x) * 4 + t] = r;
  }
}

// Position-step partials of G frames from the object patches op and their row derivatives dx
// (written by the exit-wave row pass).  Returns the chunks per frame, or 0 when m has no
// one-dimensional form (the caller then takes the 2D one).
static int pos_1d(const float2* op, const float2* dx, const float2* vp, const float2* chi, int G, int m,
                  const float2* tw, double* part, hipStream_t s) {
  auto go = [&](auto cols_kern, int nc) {
    hipLaunchKernelGGL(cols_kern, dim3(m / nc, G), dim3(PTYK_NT), 0, s, op, dx, vp, chi, tw, part);
    return m / nc;
  };
#define PTYK_POS_CASE(MM) \
  case MM: return go(k_pos_cols<MM>, (plan_min_batch<MM, PTYK_NT>() + 15) / 16 * 16);
  switch (m) {
    PTYK_POS_CASE(128)
    PTYK_POS_CASE(256)
    PTYK_POS_CASE(512)
    default: return 0;
  }
#undef PTYK_POS_CASE
}
static bool has_pos_1d(int m) { return m == 128 || m == 256 || m == 512; }

// Step clipped to 0.2 px per axis; the column-derivative step is added to the row
// coordinate and the row-derivative step to the column coordinate, as the reference does
// (probe_positions[i].x += shift from dX, Probe.cpp:1022-1033).
__global__ __launch_bounds__(1024) void k_pos_update(const double* __restrict__ part, int ch,
                                                     const int* __restrict__ gf, int G, float2* __restrict__ pos,
                                                     double* __restrict__ ftot) {
  frame_totals<1024, 4>(part, ch, G, ftot);
  if (threadIdx.x >= 64) return;  // one wave applies the steps
  for (int j = threadIdx.x; j < G; j += 64) {
    const double* s = ftot + (long)j * 4;
    const double dx = s[1] / s[0], dy = s[3] / s[2];
    const double sx = fmin(fabs(dx), 0.2) * (double)((0.0 < dx) - (dx < 0.0));
    const double sy = fmin(fabs(dy), 0.2) * (double)((0.0 < dy) - (dy < 0.0));
    const int f = gf[j];
    pos[f] = make_float2((float)((double)pos[f].x + sx), (float)((double)pos[f].y + sy));
  }
}

// Variable probe (Probe::update_variable_probe, Probe.cpp:1366-1464).
// img_j = (p / max p)_roi_j (probe_update_j - probe_update_m); partial sums of Re(V conj img_j).
__global__ __launch_bounds__(NT) void k_vp_resid(float2* __restrict__ img, const float2* __restrict__ pu,
                                                 const float2* __restrict__ pum, const float* __restrict__ illum,
                                                 const int2* __restrict__ roi, const int* __restrict__ gf, int m,
                                                 int cols, const float2* __restrict__ vmode,
                                                 const double* __restrict__ gscal, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int mm = m * m, ppb = ppb_of(m);
  const int j = blockIdx.y;
  const int2 o = roi[gf[j]];
  const float wd = (float)(1.0 / gscal[MLS_S_PMAX]);
  double s = 0.0;
  for (int t = threadIdx.x; t < ppb; t += NT) {
    const int q = blockIdx.x * ppb + t;
    const long i = (long)j * mm + q;
    const float w = illum[(long)(o.x + q / m) * cols + o.y + q % m] * wd;
    const float2 d = cscale(csub(pu[i], pum[q]), w);
    img[i] = d;
    const float2 v = vmode[q];
    s += (double)v.x * d.x + (double)v.y * d.y;
  }
  s = block_sum<NT>(s, red);
  if (threadIdx.x == 0) part[(long)j * gridDim.x + blockIdx.x] = s;
}

// c_j = (mean_px Re(V conj img_j) + e1_j) / sum_g e1_g^2  -> gscal[COEF + j]
__global__ __launch_bounds__(1024) void k_vp_coef(const double* __restrict__ part, int ch, const int* __restrict__ gf,
                                                  int G, int mm, const double* __restrict__ evo, int n,
                                                  double* __restrict__ ftot, double* __restrict__ gscal) {
  frame_totals<1024, 1>(part, ch, G, ftot);
  if (threadIdx.x >= 64) return;  // one wave does the rest (wave reductions, no block barrier)
  const int ln = threadIdx.x;
  double e = 0.0;
  for (int j = ln; j < G; j += 64) e += evo[n + gf[j]] * evo[n + gf[j]];
  e = wave_sum(e);
  for (int j = ln; j < G; j += 64) gscal[MLS_S_COEF + j] = (ftot[j] / mm + evo[n + gf[j]]) / e;
}

// acc = (1/M^2) sum_j c_j img_j; partial sums of |acc|^2 per block.
__global__ __launch_bounds__(NT) void k_vp_acc(float2* __restrict__ acc, const float2* __restrict__ img, int G, int mm,
                                               const double* __restrict__ gscal, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int q = blockIdx.x * NT + threadIdx.x;
  double s = 0.0;
  if (q < mm) {
    float2 a = make_float2(0.f, 0.f);
    int j = 0;
    constexpr int B = 8;  // eight frames' loads in flight before their sums (in frame order)
    for (; j + B <= G; j += B) {
      float2 u[B];
#pragma unroll
      for (int b = 0; b < B; ++b) u[b] = img[(long)(j + b) * mm + q];
#pragma unroll
      for (int b = 0; b < B; ++b) a = cadd(a, cscale(u[b], (float)gscal[MLS_S_COEF + j + b]));
    }
    for (; j < G; ++j) a = cadd(a, cscale(img[(long)j * mm + q], (float)gscal[MLS_S_COEF + j]));
    a = cscale(a, (float)(1.0 / mm));
    acc[q] = a;
    s = (double)a.x * a.x + (double)a.y * a.y;
  }
  s = block_sum<NT>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// V += acc * relax_U / max(1, rms(acc)); partial sums of |V|^2.
__global__ __launch_bounds__(NT) void k_vp_apply(float2* __restrict__ vmode, const float2* __restrict__ acc, int mm,
                                                 const double* __restrict__ part_in, int nb, double relax_u,
                                                 double* __restrict__ part_out) {
  __shared__ double red[NT / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) s += part_in[i];
  s = block_sum<NT>(s, red);
  const float rms = sqrtf((float)(s / mm));
  const float sc = (float)(relax_u / fmax(1.0, (double)rms));
  const int q = blockIdx.x * NT + threadIdx.x;
  double v2 = 0.0;
  if (q < mm) {
    const float2 v = cadd(vmode[q], cscale(acc[q], sc));
    vmode[q] = v;
    v2 = (double)v.x * v.x + (double)v.y * v.y;
  }
  v2 = block_sum<NT>(v2, red);
  if (threadIdx.x == 0) part_out[blockIdx.x] = v2;
}

__global__ __launch_bounds__(NT) void k_vp_norm(float2* __restrict__ vmode, int mm, const double* __restrict__ part,
                                                int nb) {
  __shared__ double red[NT / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) s += part[i];
  s = block_sum<NT>(s, red);
  const float rms = sqrtf((float)(s / mm));
  const float inv = (float)(1.0 / (double)rms);
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < mm) vmode[q] = cscale(vmode[q], inv);
}

// [sum |op_j page|^2, sum Re(chi_j conj(op_j page))] per frame chunk (d_multiplyAbsConjuReal);
// with page2, the same pair for page2 follows (four values per chunk), so the variable-probe and
// variable-intensity dots share one pass over op and chi.
__global__ __launch_bounds__(NT) void k_frame_dots(const float2* __restrict__ op, const float2* __restrict__ page,
                                                   const float2* __restrict__ page2, const float2* __restrict__ chi,
                                                   int m, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int mm = m * m, ppb = ppb_of(m);
  const int j = blockIdx.y;
  const int nv = page2 ? 4 : 2;
  double d = 0.0, u = 0.0, d2 = 0.0, u2 = 0.0;
  for (int t = threadIdx.x; t < ppb; t += NT) {
    const int q = blockIdx.x * ppb + t;
    const long i = (long)j * mm + q;
    const float2 a = op[i], b = page[q], x = chi[i];
    const double r = (double)a.x * b.x - (double)a.y * b.y, im = (double)a.x * b.y + (double)a.y * b.x;
    d += r * r + im * im;
    u += x.x * r + x.y * im;
    if (page2) {
      const float2 b2 = page2[q];
      const double r2 = (double)a.x * b2.x - (double)a.y * b2.y, im2 = (double)a.x * b2.y + (double)a.y * b2.x;
      d2 += r2 * r2 + im2 * im2;
      u2 += x.x * r2 + x.y * im2;
    }
  }
  d = block_sum<NT>(d, red);
  u = block_sum<NT>(u, red);
  double* out = part + ((long)j * gridDim.x + blockIdx.x) * nv;
  if (threadIdx.x == 0) {
    out[0] = d;
    out[1] = u;
  }
  if (page2) {
    d2 = block_sum<NT>(d2, red);
    u2 = block_sum<NT>(u2, red);
    if (threadIdx.x == 0) {
      out[2] = d2;
      out[3] = u2;
    }
  }
}

// Evolution updates: e1 += mean(u) / (mean(d) + 0.1 mean_g mean(d)) (variable probe);
// e0 += 0.1 sum(u) / sum(d) (variable intensity, Probe.cpp:1496-1516). which: 1 the first,
// 0 the second (two values per chunk), 2 both (four values per chunk, variable probe first).
__global__ __launch_bounds__(1024) void k_evo_update(const double* __restrict__ part, int ch,
                                                     const int* __restrict__ gf, int G, int mm,
                                                     double* __restrict__ evo, int n, int which,
                                                     double* __restrict__ ftot) {
  const int nv = which == 2 ? 4 : 2;
  if (nv == 4) frame_totals<1024, 4>(part, ch, G, ftot);
  else frame_totals<1024, 2>(part, ch, G, ftot);
  if (threadIdx.x >= 64) return;  // one wave does the rest (wave reductions, no block barrier)
  const int ln = threadIdx.x;
  if (which >= 1) {
    double dm = 0.0;
    for (int j = ln; j < G; j += 64) dm += ftot[nv * j] / mm;
    const double m3 = wave_sum(dm) / G;
    for (int j = ln; j < G; j += 64) evo[n + gf[j]] += (ftot[nv * j + 1] / mm) / (ftot[nv * j] / mm + 0.1 * m3);
  }
  if (which != 1) {
    const int o = which == 2 ? 2 : 0;
    for (int j = ln; j < G; j += 64) evo[gf[j]] += 0.1 * ftot[nv * j + o + 1] / ftot[nv * j + o];
  }
}

// Probe::ortho_modes (Probe.cpp:244-294): modified Gram-Schmidt in mode order with
// proj = sum(P_i conj P_j) / sum |P_j|^2, then modes 1..K-1 take the orthogonalised modes in
// order of decreasing power measured before each mode's own orthogonalisation (mode 0 stays).
// Spread over ORTHO_NB workgroups, one launch per projection (i, j): every workgroup finishes the
// coefficient from the previous launch's block partials (the same fixed-order sum in each),
// subtracts the projection on its pixels and leaves the partials of the next pair (ni, nj) from
// the values it has just written.  A pixel belongs to the same thread in every launch.  The first
// launch (i < 0) leaves the power partials of every mode (mode i's sweep changes only mode i, so
// they are the powers before each sweep); the last (ni < 0) also permutes its pixels' modes.
constexpr int ORTHO_NB = 64;
constexpr int ORTHO_KMAX = 16;
__global__ __launch_bounds__(NT) void k_ortho_step(float2* __restrict__ probe, int mm, int K, int i, int j, int ni,
                                                   int nj, const double* __restrict__ pin, double* __restrict__ pout,
                                                   double* __restrict__ pw) {
  __shared__ double red[NT / 64];
  __shared__ double coef[2];
  __shared__ int order[ORTHO_KMAX];
  const int stride = gridDim.x * NT;
  if (i >= 0) {
    if (threadIdx.x < 64) {
      double a = 0.0, b = 0.0, c = 0.0;
      for (int t = threadIdx.x; t < (int)gridDim.x; t += 64) {
        a += pin[3 * t];
        b += pin[3 * t + 1];
        c += pin[3 * t + 2];
      }
      a = wave_sum(a);
      b = wave_sum(b);
      c = wave_sum(c);
      if (threadIdx.x == 0) {
        coef[0] = b / a;
        coef[1] = c / a;
      }
    }
    __syncthreads();
    const float2 pr = c_of(coef[0], coef[1]);
    float2* pi = probe + (long)i * mm;
    const float2* pj = probe + (long)j * mm;
    for (int q = blockIdx.x * NT + threadIdx.x; q < mm; q += stride) pi[q] = csub(pi[q], cmul(pj[q], pr));
  } else {
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int q = blockIdx.x * NT + threadIdx.x; q < mm; q += stride) s += cabs2(probe[(long)k * mm + q]);
      s = block_sum<NT>(s, red);
      if (threadIdx.x == 0) pw[(long)blockIdx.x * K + k] = s;
    }
  }
  if (ni >= 0) {
    const float2* a = probe + (long)ni * mm;
    const float2* b = probe + (long)nj * mm;
    double sr = 0.0, slr = 0.0, sli = 0.0;
    for (int q = blockIdx.x * NT + threadIdx.x; q < mm; q += stride) {
      const float2 u = a[q], v = b[q];
      sr += (double)v.x * v.x + (double)v.y * v.y;
      slr += (double)u.x * v.x + (double)u.y * v.y;
      sli += (double)u.y * v.x - (double)u.x * v.y;
    }
    sr = block_sum<NT>(sr, red);
    slr = block_sum<NT>(slr, red);
    sli = block_sum<NT>(sli, red);
    if (threadIdx.x == 0) {
      pout[3 * blockIdx.x] = sr;
      pout[3 * blockIdx.x + 1] = slr;
      pout[3 * blockIdx.x + 2] = sli;
    }
    return;
  }
  // last launch: order modes 1..K-1 by (-power, index) and permute this thread's pixels
  if (threadIdx.x == 0) {
    double pk[ORTHO_KMAX];
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int b = 0; b < (int)gridDim.x; ++b) t += pw[(long)b * K + k];
      pk[k] = t;
      order[k] = k;
    }
    for (int a = 1; a < K; ++a)  // insertion sort, stable
      for (int b = a; b > 0 && pk[order[b]] > pk[order[b - 1]]; --b) {
        const int t = order[b];
        order[b] = order[b - 1];
        order[b - 1] = t;
      }
  }
  __syncthreads();
  for (int q = blockIdx.x * NT + threadIdx.x; q < mm; q += stride) {
    float2 v[ORTHO_KMAX];
#pragma unroll
    for (int k = 1; k < ORTHO_KMAX; ++k)
      if (k < K) v[k] = probe[(long)k * mm + q];
#pragma unroll
    for (int k = 1; k < ORTHO_KMAX; ++k)
      if (k < K) {
        const int src = order[k];
        // mode 0 keeps its place; a mode >= 1 moves to slot k (order[0] is 0: power sorts
        // include mode 0 but slot 0 is never rewritten, as in the reference's copy loop)
        probe[(long)k * mm + q] = src == 0 ? probe[q] : v[src];
      }
  }
}

inline dim3 g1(long count) { return dim3((unsigned)((count + NT - 1) / NT)); }

// First kernel of a group graph: this group's frames (sel[1 ..]) and the next group's
// (sel[1 + gmax ..]) from the iteration's table at the group index sel[0], which then advances.
__global__ void k_select_group(const int* __restrict__ gframes, const int* __restrict__ goffs, int has_next,
                               int gmax, int* __restrict__ sel) {
  const int g = sel[0];
  const int a = goffs[g], b = goffs[g + 1];
  for (int i = threadIdx.x; i < b - a; i += blockDim.x) sel[1 + i] = gframes[a + i];
  if (has_next) {
    const int c = goffs[g + 2];
    for (int i = threadIdx.x; i < c - b; i += blockDim.x) sel[1 + gmax + i] = gframes[b + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) sel[0] = g + 1;
}

__global__ void k_set_int(int* __restrict__ p, int v) { *p = v; }

}  // namespace

int mls_chunks(int m) { return m * m / ppb_of(m); }

#define LAUNCH(kernel, grid, block, ...) hipLaunchKernelGGL(kernel, grid, block, 0, s, __VA_ARGS__)
#define CHK()                               \
  do {                                      \
    hipError_t e_ = hipGetLastError();      \
    if (e_ != hipSuccess) return e_;        \
  } while (0)
#define TRY(x)                              \
  do {                                      \
    hipError_t e_ = (x);                    \
    if (e_ != hipSuccess) return e_;        \
  } while (0)

// Probe::ortho_modes as 1 + K (K - 1) / 2 launches of k_ortho_step (partials ping-pong in d.part).
hipError_t mls_ortho(MlsDev& d, hipStream_t s) {
  const int K = d.k, mm = d.m * d.m;
  if (K < 2) return hipSuccess;
  if (K > ORTHO_KMAX || d.part_cap < 6L * ORTHO_NB + (long)K * ORTHO_NB) return hipErrorInvalidValue;
  double* ping = d.part;
  double* pong = d.part + 3 * ORTHO_NB;
  double* pw = d.part + 6 * ORTHO_NB;
  // (1, 0), (2, 0), (2, 1), ..., (K - 1, K - 2); the first launch only prepares (1, 0)
  std::vector<std::pair<int, int>> pairs;
  for (int i = 1; i < K; ++i)
    for (int j = 0; j < i; ++j) pairs.emplace_back(i, j);
  for (size_t t = 0; t <= pairs.size(); ++t) {
    const int ci = t == 0 ? -1 : pairs[t - 1].first, cj = t == 0 ? -1 : pairs[t - 1].second;
    const int ni = t < pairs.size() ? pairs[t].first : -1, nj = t < pairs.size() ? pairs[t].second : -1;
    LAUNCH(k_ortho_step, dim3(ORTHO_NB), dim3(NT), d.probe, mm, K, ci, cj, ni, nj, ping, pong, pw);
    CHK();
    std::swap(ping, pong);
  }
  return hipSuccess;
}

void mls_drop_graphs(MlsDev& d) {
  for (auto& kv : d.graphs) hipGraphExecDestroy(kv.second);
  d.graphs.clear();
}

hipError_t mls_roi(MlsDev& d, hipStream_t s) {
  const float cx = (float)std::ceil(d.rows * 1.0 / 2 - d.m * 1.0 / 2);
  const float cy = (float)std::ceil(d.cols * 1.0 / 2 - d.m * 1.0 / 2);
  LAUNCH(k_roi, g1(d.n), dim3(NT), d.pos, d.n, d.m, d.rows, d.cols, cx, cy, d.roi, d.sub, d.flag);
  return hipGetLastError();
}

// Canvas / array reduction finished on the host (initialisation only).
static hipError_t host_reduce(MlsDev& d, const float* a, const float2* o, long count, int op, double* out,
                              hipStream_t s) {
  const int nb = (int)std::min<long>((count + NT - 1) / NT, 1024);
  LAUNCH(k_reduce, dim3(nb), dim3(NT), a, o, count, op, d.part);
  CHK();
  std::vector<double> h(nb);
  TRY(hipMemcpyAsync(h.data(), d.part, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  double r = 0.0;
  for (int i = 0; i < nb; ++i) r = (op == 2) ? std::max(r, h[i]) : r + h[i];
  *out = r;
  return hipSuccess;
}

hipError_t mls_init(MlsDev& d, const double* frame_power, hipStream_t s) {
  const int mm = d.m * d.m;
  const long canvas = (long)d.rows * d.cols;
  std::vector<double> evo(2 * (size_t)d.n);
  TRY(hipMemcpyAsync(evo.data(), d.evo, evo.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  for (int i = 0; i < d.n; ++i) evo[i] = 1.0;
  // variable mode to unit rms, its evolution scaled by the same norm (h_norm2: float result)
  if (d.variable_probe_modes > 0) {
    TRY(sum_abs2(d.vmode, mm, d.gscal + MLS_S_TMP, s));
    double sv = 0.0;
    TRY(hipMemcpyAsync(&sv, d.gscal + MLS_S_TMP, sizeof(double), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    const float rv = sqrtf((float)(sv / mm));
    const double inv = (float)(1.0 / (double)rv);
    TRY(hipMemcpyAsync(d.gscal + MLS_S_TMP, &inv, sizeof(double), hipMemcpyHostToDevice, s));
    LAUNCH(k_scale_c, g1(mm), dim3(NT), d.vmode, (long)mm, d.gscal + MLS_S_TMP, 0);
    CHK();
    for (int i = 0; i < d.n; ++i) evo[d.n + i] *= rv;
  }
  TRY(hipMemcpyAsync(d.evo, evo.data(), evo.size() * sizeof(double), hipMemcpyHostToDevice, s));
  // illumination = sum of |P0|^2 at every ROI
  LAUNCH(k_illum, dim3((d.cols + 63) / 64, (d.rows + 3) / 4), dim3(NT), d.illum, d.probe, d.roi, d.n, d.m, d.rows,
         d.cols, 0);
  CHK();
  // object normalised by the illumination-weighted rms
  double si = 0.0, soi = 0.0;
  TRY(host_reduce(d, d.illum, nullptr, canvas, 0, &si, s));
  TRY(host_reduce(d, d.illum, d.obj, canvas, 1, &soi, s));
  const float onorm = (float)std::sqrt(soi / si);
  double f = (float)(1.0 / (double)onorm);
  TRY(hipMemcpyAsync(d.gscal + MLS_S_TMP, &f, sizeof(double), hipMemcpyHostToDevice, s));
  LAUNCH(k_scale_c, g1(canvas), dim3(NT), d.obj, canvas, d.gscal + MLS_S_TMP, 0);
  CHK();
  // probe power matched to the data: corr = sqrt(median_j(sum|P|^2 / sum I_j) M^2)
  TRY(sum_abs2(d.probe, (long)d.k * mm, d.gscal + MLS_S_TMP + 1, s));
  double pw = 0.0;
  TRY(hipMemcpyAsync(&pw, d.gscal + MLS_S_TMP + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  const double meanillum = pw / mm;
  std::vector<double> cv(d.n);
  for (int i = 0; i < d.n; ++i) cv[i] = meanillum / (frame_power[i] / mm);
  std::sort(cv.begin(), cv.end());
  const double med = (d.n % 2 == 0) ? (cv[d.n / 2 - 1] + cv[d.n / 2]) / 2 : cv[d.n / 2];
  const float corr = (float)std::sqrt(med * mm);
  f = (float)(1.0 / (double)corr);
  TRY(hipMemcpyAsync(d.gscal + MLS_S_TMP, &f, sizeof(double), hipMemcpyHostToDevice, s));
  LAUNCH(k_scale_c, g1((long)d.k * mm), dim3(NT), d.probe, (long)d.k * mm, d.gscal + MLS_S_TMP, 0);
  LAUNCH(k_scale_c, g1(mm), dim3(NT), d.vmode, (long)mm, d.gscal + MLS_S_TMP, 0);
  CHK();
  // illum /= corr^2 (float product), MAX_ILLUM = max illum
  {
    std::vector<float> h(canvas);
    TRY(hipMemcpyAsync(h.data(), d.illum, canvas * sizeof(float), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    const float c2 = corr * corr;
    float mx = 0.f;
    for (long i = 0; i < canvas; ++i) {
      h[i] = h[i] / c2;
      mx = std::max(mx, h[i]);
    }
    TRY(hipMemcpyAsync(d.illum, h.data(), canvas * sizeof(float), hipMemcpyHostToDevice, s));
    TRY(hipStreamSynchronize(s));
    d.max_illum = mx;
  }
  // LSQ steps start at 1 (IPhaser.cpp:343-349)
  LAUNCH(k_fill_c, g1(d.n), dim3(NT), d.step_obj, (long)d.n, make_float2(1.f, 0.f));
  LAUNCH(k_fill_c, g1(d.n), dim3(NT), d.step_probe, (long)d.n, make_float2(1.f, 0.f));
  CHK();
  return hipStreamSynchronize(s);
}

long mls_part_min(int k) { return (6L + std::max(k, ORTHO_KMAX)) * ORTHO_NB; }

SideEnqueuer::SideEnqueuer(int device) : device_(device), th_([this] { run(); }) {}

SideEnqueuer::~SideEnqueuer() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_one();
  th_.join();
}

void SideEnqueuer::post(std::function<hipError_t()> task) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    q_.push_back(std::move(task));
    ++posted_;
  }
  cv_.notify_one();
}

hipError_t SideEnqueuer::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [this] { return done_ == posted_; });
  const hipError_t e = err_;
  err_ = hipSuccess;
  return e;
}

void SideEnqueuer::run() {
  const hipError_t set = hipSetDevice(device_);  // the current device is per host thread
  for (;;) {
    std::function<hipError_t()> task;
    bool skip;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
      if (q_.empty()) return;  // stop_ with nothing left
      task = std::move(q_.front());
      q_.pop_front();
      // after an error the remaining tasks are dropped (their waits and launches would build on
      // a schedule that is already broken); drain() reports the first error
      skip = err_ != hipSuccess;
    }
    const hipError_t e = set != hipSuccess ? set : (skip ? hipSuccess : task());
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (e != hipSuccess && err_ == hipSuccess) err_ = e;
      ++done_;
    }
    cv_done_.notify_all();
  }
}

#define MARK(c) \
  do {          \
    if (tm) tm->mark(s, (c)); \
  } while (0)

static hipError_t mls_iteration_body(MlsDev& d, int iter, const int* goffs, int n_groups, hipStream_t s,
                                     MlsTimer* tm) {
  const int m = d.m, mm = m * m, K = d.k, ch = mls_chunks(m);
  const long canvas = (long)d.rows * d.cols;
  const dim3 tiles((d.cols + 63) / 64, (d.rows + 3) / 4);
  const dim3 pre_tiles((d.cols + 63) / 64, (d.rows + 4 * PRE_RPT - 1) / (4 * PRE_RPT));
  const bool var_any = d.variable_probe || d.variable_intensity;
  const int subpix = d.apply_subpix_shift ? 1 : 0;

  MARK(MLS_T_ITER);
  // MLS.cpp:86-89
  if (iter == 1 || iter >= d.pps) TRY(mls_roi(d, s));
  // calc_object_norm over the precalculated box
  {
    const int x0 = d.box[0] - 1, y0 = d.box[2] - 1;
    const int bx = d.box[1] - d.box[0] + 1, by = d.box[3] - d.box[2] + 1;
    LAUNCH(k_box_partial, dim3(bx), dim3(NT), d.illum, d.obj, d.cols, x0, y0, by, d.part);
    LAUNCH(k_objnorm, dim3(1), dim3(NT), d.part, bx, d.gscal);
    LAUNCH(k_scale_c, g1((long)K * mm), dim3(NT), d.probe, (long)K * mm, d.gscal + MLS_S_ONORM, 0);
    LAUNCH(k_scale_c, g1(mm), dim3(NT), d.vmode, (long)mm, d.gscal + MLS_S_ONORM, 0);
    LAUNCH(k_scale_c, g1(canvas), dim3(NT), d.obj, canvas, d.gscal + MLS_S_ONORM, 1);
    CHK();
  }
  LAUNCH(k_pos_diff, dim3(1), dim3(NT), d.pos, d.pos_o, d.n);
  if (d.variable_probe_modes && iter > d.probe_reconstruct) {
    TRY(sum_abs2(d.vmode, mm, d.gscal + MLS_S_TMP, s));
    LAUNCH(k_remove_extra, dim3(1), dim3(NT), d.vmode, d.probe, mm, d.evo, d.n, d.gscal);
  }
  LAUNCH(k_illum, tiles, dim3(NT), d.illum, d.probe, d.roi, d.n, m, d.rows, d.cols, 1);
  CHK();
  const bool var_update = var_any && iter > d.probe_reconstruct + 1;
  if (var_update && d.variable_probe) {  // max p is fixed for the rest of the iteration
    const int nb = (int)std::min<long>((canvas + NT - 1) / NT, 1024);
    LAUNCH(k_reduce, dim3(nb), dim3(NT), d.illum, nullptr, canvas, 2, d.part);
    LAUNCH(k_finish_max, dim3(1), dim3(NT), d.part, nb, d.gscal + MLS_S_PMAX);
    CHK();
  }
  const float fw = (float)((double)d.max_illum * 0.01);
  const float fd = (float)((double)d.max_illum * d.delta_p);
  const bool obj_rec_iter = iter >= d.object_reconstruct;
  const bool probe_rec = iter >= d.probe_reconstruct;

  const int mod_blocks = mls_modulus_blocks(m, K, d.mod_cb);
  // the illumination probes of group g + 1 (spectra of P_0 and V, shifted probes) are formed on
  // the second stream while the probe modes 2..K of group g are updated: they read only P_0, V,
  // the evolution coefficients and sub-pixel shifts of group g + 1's frames, none of which the
  // rest of group g writes, and they write only spec and vp, which nothing after that point in
  // group g reads
  const bool prefetch = subpix && mod_blocks && d.side;
  // Second-stream work of a group, on two streams: A takes the position step and then the next
  // group's illumination probes (which overwrite vp, read by the position step), B the
  // variable-mode update and then the evolution dots (which read the updated V).  Each piece waits
  // on its own fork event, recorded on s where the piece's inputs are final, and the illumination
  // probes also on the end of B's variable-mode update (they read V).  The pieces are enqueued
  // after the group's main-stream launches, so the host submits those first.
  // Pieces whose inputs become final at the same point share one fork event (each event record is
  // a marker between two main-stream kernels, a few microseconds of the main stream each).
  struct SideItem {
    int lane;               // 0: stream A, 1: stream B
    int point;              // its fork event, ev_side[point]
    bool after_vp;          // also waits for the variable-mode update of this group
    bool signal_vp;         // this is the variable-mode update: its end is recorded
    std::function<hipError_t(hipStream_t)> work;
  };
  SideItem side_work[MLS_SIDE_FORKS];
  int n_side = 0, n_points = 0;
  auto fork_point = [&]() -> hipError_t {
    if (n_points >= MLS_SIDE_FORKS) return hipErrorInvalidValue;
    return hipEventRecord(d.ev_side[n_points++], s);
  };
  // The join of a group's side streams can be left to the start of the next group (not under
  // graph capture, where every group's capture must end joined): join_lane is the stream whose
  // end event the main stream still has to wait on, join_G the size of that group.  With the
  // join deferred, a second host thread (d.enq) enqueues each side piece as soon as it forks,
  // while this thread goes on with the main stream; the main stream's wait on the join event
  // follows the drain of that thread (so the event has been recorded), and so does every fork
  // event's next record (the side pieces' waits on it have been enqueued).
  bool defer_join = false;
  int join_lane = -1, join_G = 0;
  bool side_used[2] = {false, false};  // streams A and B took work this group
  auto side_task = [&d](const SideItem& w) {
    hipStream_t st = w.lane ? d.side2 : d.side;
    return [&d, st, point = w.point, after_vp = w.after_vp, signal_vp = w.signal_vp,
            work = w.work]() -> hipError_t {
      TRY(hipStreamWaitEvent(st, d.ev_side[point], 0));
      if (after_vp) TRY(hipStreamWaitEvent(st, d.ev_vpdone, 0));
      TRY(work(st));
      if (signal_vp) TRY(hipEventRecord(d.ev_vpdone, st));
      return hipSuccess;
    };
  };
  auto fork = [&](int lane, bool after_vp, bool signal_vp, std::function<hipError_t(hipStream_t)> work) -> hipError_t {
    if (n_side >= MLS_SIDE_FORKS || n_points == 0) return hipErrorInvalidValue;
    SideItem w{lane, n_points - 1, after_vp, signal_vp, std::move(work)};
    side_used[lane] = true;
    if (defer_join && d.enq) {
      d.enq->post(side_task(w));  // enqueued now, by the second host thread
    } else {
      side_work[n_side++] = std::move(w);  // enqueued after the group's main-stream launches
    }
    return hipSuccess;
  };
  // the deferred join: the second host thread's tasks have all run (the join event is recorded)
  auto join_wait = [&]() -> hipError_t {
    if (d.enq) TRY(d.enq->drain());
    TRY(hipStreamWaitEvent(s, d.ev_join[join_lane], 0));
    join_lane = -1;
    return hipSuccess;
  };
  // One group's work on both streams; gf and gfn are the frame lists of this group and the next.
  auto group_body = [&](int g, const int* gf, const int* gfn) -> hipError_t {
    const int G = goffs[g + 1] - goffs[g];
    const bool vp_ready = prefetch && g > 0;  // formed by the previous group's second-stream work
    // object patches and dX rows alternate between two buffers from group to group: this group's
    // are formed while the previous group's side work still reads that group's
    float2* const opg = (g & 1) ? d.op_alt : d.op;
    float2* const dxg = (g & 1) ? d.dxb_alt : d.dxb;
    bool vp_on_side = false;                  // the variable-mode update runs on stream B this group
    // side pieces waiting for their fork point
    std::function<hipError_t(hipStream_t)> pos_work, vp_work, evo_work;
    const long gm = (long)G * mm;
    const dim3 gpix((mm + NT - 1) / NT, G), gch(ch, G);
    // illumination probe, exit waves, Fourier transforms
    const bool pos_step = iter >= d.pps;
    // with the three-pass exit waves the object patches stay in op, and the position step's
    // derivative images come from one-dimensional transforms of them
    const bool pos_1d_step = pos_step && mod_blocks && has_pos_1d(m) && d.side;
    const bool spec2d = pos_step && !pos_1d_step;
    // the exit waves of modes 1..K-1 read only O and P_1..K-1 (final on this stream) and write
    // only their planes G .. K G - 1, past every plane the previous group's side work reads (its
    // chi_0 planes 0 .. join_G - 1) when G >= join_G: they go ahead of the join, so the main
    // stream has work while the previous group's side streams finish
    bool modes_early = false;
    bool point_bound = false;   // the probe-update average's launch recorded the first fork event
    bool point2_bound = false;  // the step application's launch recorded the second
    const MlsExitArgs ex0{d.obj, d.cols, d.probe, d.vp, d.roi, gf, subpix, opg, spec2d ? d.tmp : nullptr,
                          pos_1d_step ? dxg : nullptr};
    if (join_lane >= 0) {
      if (vp_ready && mod_blocks && d.row_waves && (m == 128 || m == 256) && K > 1 && G >= join_G && !spec2d) {
        // with mode 0's object patches and dX rows (into this group's buffers; its exit waves wait
        // for the illumination probes)
        MlsExitArgs ex_early = ex0;
        ex_early.mode0 = 1;
        MARK(MLS_T_PASS_A);
        TRY(mls_exit_rows(d.psi, G, m, &ex_early, 0, K * G, d.tw, s, tm));
        modes_early = true;
      }
      MARK(MLS_T_JOIN);  // the wait is charged to the join, not to the launch before it
      TRY(join_wait());
    }
    MARK(MLS_T_PROBES);
    if (subpix && !vp_ready && !mod_blocks) TRY(fft2_pair(d.spec, d.probe, d.vmode, m, d.tw, s));  // spectra of P_0 and V
    // modes 2..K only feed the fused probe-mode update: their inverse row pass is left to it
    const bool defer_modes = probe_rec && !d.apply_multimodal_update && K > 1 && mod_blocks && d.pum_part;
    const bool pu_fused = defer_modes && subpix;
    if (subpix && mod_blocks) {
      // the shifted illumination probes from the row spectra of P_0 and V, column pass first
      if (!vp_ready) {
        TRY(fft_rows_pair(d.spec, d.probe, d.vmode, m, d.tw, s));
        TRY(vp_from_row_spectra(d.vp, d.spec, d.evo, d.sub, gf, d.n, var_any ? 1 : 0, G, m, d.tw, s));
      }
    } else {
      LAUNCH(k_vp, gpix, dim3(NT), d.vp, d.spec, d.probe, d.vmode, d.evo, d.sub, gf, d.n, m, subpix, var_any ? 1 : 0);
      CHK();
      if (subpix) TRY(fft2_batch(d.vp, G, m, true, d.tw, s));
    }
    if (mod_blocks) {
      // three passes for the exit waves, the first of which forms them (and the object patches)
      // from O, P and vp; the position step's patch copy takes its own forward transform; with
      // the one-dimensional position step, mode 0's exit-wave workgroups also form its dX rows
      MlsExitArgs ex = ex0;
      if (modes_early) {  // mode 0's exit waves only, from the patches formed ahead of the join
        ex.np = G;
        ex.mode0 = 2;
      }
      // with mode 0 alone in the inverse row pass, that pass also forms the probe update's row shifts
      const MlsShiftArgs shf{d.pu, opg, d.sub, -1.f};
      TRY(mls_modulus_passes(d.psi, K, G, m, d.mod_cb, d.row_waves, d.amp, d.amp_scale, gf, d.part, d.tw, &ex,
                             defer_modes ? 1 : K, pu_fused ? &shf : nullptr, s, tm));
      if (spec2d) TRY(fft2_batch(d.tmp, G, m, false, d.tw, s));
      if (pos_1d_step) {
        // the position step reads only chi0, vp and op, which stay fixed until the group ends,
        // and writes only its own buffers and the positions: it runs beside the rest of the
        // group (forked here rather than after the step application, which measured 0.8 ms per
        // 51x51 / 256^2 / 5-mode iteration slower)
        pos_work = [&d, G, m, gf, opg, dxg](hipStream_t st) -> hipError_t {
          // dX came with the exit waves (k_mls_pass_a)
          const int pch = pos_1d(opg, dxg, d.vp, d.psi, G, m, d.tw, d.pos_part, st);
          hipLaunchKernelGGL(k_pos_update, dim3(1), dim3(1024), 0, st, d.pos_part, pch, gf, G, d.pos, d.pos_ftot);
          return hipGetLastError();
        };
      }
    } else {
      LAUNCH(k_psi, gpix, dim3(NT), d.psi, opg, pos_step ? d.tmp : nullptr, d.vp, d.obj, d.probe, d.roi, gf, G, K,
             m, d.cols, subpix);
      CHK();
      // with the position step, tmp (the object patch, right after the exit-wave buffer) rides
      // in the same forward transform (frames between K G and K gmax are unused scratch)
      TRY(fft2_batch(d.psi, pos_step ? (long)K * d.gmax + G : (long)K * G, m, false, d.tw, s));
      LAUNCH(k_modulus, gch, dim3(NT), d.psi, d.amp, d.amp_scale, gf, G, K, m, d.part);
      CHK();
      TRY(fft2_batch(d.psi, (long)K * G, m, true, d.tw, s));
    }
    MARK(MLS_T_PROBE_UPDATE);
    if (iter < 10) LAUNCH(k_ferr, g1(G), dim3(NT), d.part, mod_blocks ? mod_blocks : ch, gf, G, mm, d.ferr);
    if (!probe_rec) {  // probe_update / probe_update_m stay zero (MLS.cpp:146-153)
      LAUNCH(k_fill_c, g1(gm), dim3(NT), d.pu, gm, make_float2(0.f, 0.f));
      LAUNCH(k_fill_c, g1(mm), dim3(NT), d.pum, (long)mm, make_float2(0.f, 0.f));
    }
    CHK();
    const float2* chi0 = d.psi;
    const int ll_range = std::max(K, d.n_objects);
    // modes 2..K take one fused pass when no object update runs for them
    const bool fuse_modes = probe_rec && !d.apply_multimodal_update && K > 1 && K <= 16;
    // the variable-mode update on the second stream; the 2D position step (which rewrites tmp,
    // the update's scratch) rules it out
    // (and with later modes' own probe-update passes, which rewrite pum, ruled out too)
    const bool vp_side = var_update && d.variable_probe && probe_rec && d.side && d.vp_part && mod_blocks &&
                         subpix && (!pos_step || pos_1d_step) && (fuse_modes || std::max(K, d.n_objects) == 1);
    for (int ll = 1; ll <= ll_range; ++ll) {
      if (fuse_modes && ll > 1) break;
      const bool obj_rec = obj_rec_iter && (d.apply_multimodal_update || ll <= d.n_objects);
      const int llp = std::min(K, ll);
      const float2* chi = d.psi + (long)(ll - 1) * gm;
      bool pre_done = false;  // the object direction already ran with the probe-update average
      if (probe_rec) {
        if (ll == 1) {
          if (subpix && mod_blocks) {
            // chi0 conj(op) with its sub-pixel shift in two passes (row shifts, column shifts); the
            // row shifts already ran inside the inverse row pass when pu_fused
            if (pu_fused)
              TRY(mls_shift_cols(d.pu, G, m, d.sub, gf, -1.f, d.tw, s));
            else
              TRY(mls_shift_passes(d.pu, chi0, opg, G, m, d.sub, gf, -1.f, d.tw, s));
          } else {
            LAUNCH(k_pu, g1(gm), dim3(NT), d.pu, chi0, opg, gm);
            CHK();
            if (subpix) {
              TRY(fft2_batch(d.pu, G, m, false, d.tw, s));
              LAUNCH(k_ramp, gpix, dim3(NT), d.pu, d.sub, gf, m, -1.f);
              CHK();
              TRY(fft2_batch(d.pu, G, m, true, d.tw, s));
            }
          }
          // with the object update, mode 1's object direction (chi0 and vp, final since the inverse
          // pass) runs in the same launch as the probe-update average
          pre_done = obj_rec;
          const int n_pum = (mm + NT - 1) / NT;
          const dim3 g_pp((unsigned)n_pum + pre_tiles.x * pre_tiles.y);
          const int scatter = d.delta_p != 0 ? 1 : 0;
          hipEvent_t stop = nullptr;
          if (defer_join && (static_cast<bool>(pos_work) || vp_side) && n_points < MLS_SIDE_FORKS) {
            // the side pieces' fork event is the end of this launch itself (recorded by its
            // dispatch, with no marker between it and the next launch on this stream)
            stop = d.ev_side[n_points++];
            point_bound = true;
          }
          if (pre_done && stop)
            hipExtLaunchKernelGGL(k_pum_precond<PRE_RPT>, g_pp, dim3(NT), 0, s, nullptr, stop, 0, d.pum,
                                  (const float2*)d.pu, fw, n_pum, d.precond, (const float*)d.illum, chi0,
                                  (const float2*)d.vp, (const float2*)nullptr, (const int2*)d.roi, gf, G, m, d.rows,
                                  d.cols, fd, scatter, (int)pre_tiles.x);
          else if (pre_done)
            LAUNCH(k_pum_precond<PRE_RPT>, g_pp, dim3(NT), d.pum, d.pu, fw, n_pum, d.precond, d.illum, chi0, d.vp,
                   (const float2*)nullptr, d.roi, gf, G, m, d.rows, d.cols, fd, scatter, (int)pre_tiles.x);
          else if (stop)
            hipExtLaunchKernelGGL(k_pum, g1(mm), dim3(NT), 0, s, nullptr, stop, 0, d.pum, (const float2*)d.pu,
                                  (const float2*)opg, (const float*)d.illum, (const int2*)d.roi, gf, G, m, d.cols, fw,
                                  0);
          else
            LAUNCH(k_pum, g1(mm), dim3(NT), d.pum, d.pu, opg, d.illum, d.roi, gf, G, m, d.cols, fw, 0);
          if (vp_side) {
            // the variable-mode update needs only pu, pum, the illumination and V, and writes
            // only V, its own scratch and the coefficients: it runs on the second stream beside
            // the LSQ steps and their application (and the object direction when that runs apart)
            CHK();
            vp_work = [&d, G, m, mm, ch, gch, gf](hipStream_t st) -> hipError_t {
              const int nb = (mm + NT - 1) / NT;
              double* p2 = d.vp_part + (long)G * ch;  // |acc|^2 partials
              double* p3 = p2 + nb;                   // |V|^2 partials
              hipLaunchKernelGGL(k_vp_resid, gch, dim3(NT), 0, st, d.tmp, d.pu, d.pum, d.illum, d.roi, gf, m, d.cols,
                                 d.vmode, d.gscal, d.vp_part);
              hipLaunchKernelGGL(k_vp_coef, dim3(1), dim3(1024), 0, st, d.vp_part, ch, gf, G, mm, d.evo, d.n,
                                 d.vp_ftot, d.gscal);
              hipLaunchKernelGGL(k_vp_acc, dim3(nb), dim3(NT), 0, st, d.acc, d.tmp, G, mm, d.gscal, p2);
              hipLaunchKernelGGL(k_vp_apply, dim3(nb), dim3(NT), 0, st, d.vmode, d.acc, mm, p2, nb, (double)G / d.n,
                                 p3);
              hipLaunchKernelGGL(k_vp_norm, dim3(nb), dim3(NT), 0, st, d.vmode, mm, p3, nb);
              return hipGetLastError();
            };
            vp_on_side = true;
          }
        } else {
          LAUNCH(k_pum, g1(mm), dim3(NT), d.pum, chi, opg, d.illum, d.roi, gf, G, m, d.cols, fw, 1);
        }
        CHK();
      }
      const float2* pl = llp == 1 ? nullptr : d.probe + (long)(llp - 1) * mm;
      if (ll == 1 && (pos_work || vp_work)) {
        // the position step (its inputs final since the inverse row pass) and the variable-mode
        // update (since the probe-update average) fork here, at one point
        if (!point_bound) TRY(fork_point());
        if (pos_work) TRY(fork(0, false, false, std::move(pos_work)));
        if (vp_work) TRY(fork(1, false, true, std::move(vp_work)));
      }
      MARK(MLS_T_PRECOND);
      if (obj_rec && !pre_done) {
        const int scatter = d.delta_p != 0 ? 1 : 0;
        LAUNCH(k_precond<PRE_RPT>, pre_tiles, dim3(NT), d.precond, d.illum, chi, d.vp, pl, d.roi, gf, G, m, d.rows, d.cols, fd,
               scatter);
        CHK();
      }
      const bool lsq = d.beta_lsq > 0 && ll == 1 && obj_rec && probe_rec;
      MARK(MLS_T_LSQ);
      if (lsq) {
        // the object direction: the preconditioned update at each ROI, or chi conj(vp) unscaled
        const bool scatter = d.delta_p != 0;
        if (!scatter) LAUNCH(k_dir, gpix, dim3(NT), d.tmp2, chi, d.vp, pl, m);
        LAUNCH(k_lsq_partial, gch, dim3(NT), d.tmp2, scatter ? d.precond : nullptr, d.roi, gf, d.cols, d.vp, d.pum,
               opg, chi0, m, d.part);
        CHK();
      }
      const double lam_mm = (double)(float)(1.1920928955078125e-07 / mm) * mm;
      if (ll == 1 && (obj_rec_iter || probe_rec)) {  // the group's steps do not change for ll > 1
        LAUNCH(k_lsq_solve, dim3(1), dim3(1024), d.part, ch, gf, G, lam_mm, d.beta_object * d.beta_lsq,
               d.beta_probe * d.beta_lsq, lsq ? 1 : 0, d.step_obj, d.step_probe, d.ftot, d.gscal);
        CHK();
      }
      MARK(MLS_T_STEP);
      {  // object and probe steps in one launch (disjoint arrays, both after the step solve)
        const long no = obj_rec ? canvas : 0, np = probe_rec && ll <= K ? (long)mm : 0;
        // the second fork point (evolution dots, next group's illumination probes) is the end of
        // this launch when nothing else of the group runs on this stream before that point
        const bool evo_side = var_update && d.variable_probe && vp_side;
        const bool bind2 = defer_join && ll == 1 && no + np > 0 && (fuse_modes || ll_range == 1) &&
                           (!pos_step || pos_1d_step) && (!var_update || evo_side) &&
                           (evo_side || (prefetch && g + 1 < n_groups)) && n_points < MLS_SIDE_FORKS;
        if (bind2) {
          hipExtLaunchKernelGGL(k_axpy_beta, g1((no + 1) / 2 + (np + 1) / 2), dim3(NT), 0, s, nullptr,
                                d.ev_side[n_points], 0, d.obj, (const float2*)d.precond, no,
                                (const double*)(d.gscal + MLS_S_BO), d.probe + (long)(ll - 1) * mm,
                                (const float2*)d.pum, np, (const double*)(d.gscal + MLS_S_BP));
          ++n_points;
          point2_bound = true;
        } else if (no + np > 0) {
          LAUNCH(k_axpy_beta, g1((no + 1) / 2 + (np + 1) / 2), dim3(NT), d.obj, d.precond, no, d.gscal + MLS_S_BO,
                 d.probe + (long)(ll - 1) * mm, d.pum, np, d.gscal + MLS_S_BP);
        }
      }
      CHK();
      if (pos_step && ll == 1 && !pos_1d_step) {
        LAUNCH(k_deriv, gpix, dim3(NT), d.tmp, d.tmp2, m);  // tmp already holds its spectrum
        CHK();
        // tmp and tmp2 in one inverse transform (tmp2 follows tmp's max(gmax, K) frames)
        TRY(fft2_batch(d.tmp, std::max<long>(d.gmax, K) + G, m, true, d.tw, s));
        LAUNCH(k_pos_partial, gch, dim3(NT), d.tmp, d.tmp2, d.vp, chi0, m, d.part);
        LAUNCH(k_pos_update, dim3(1), dim3(1024), d.part, ch, gf, G, d.pos, d.ftot);
        CHK();
      }
      if (var_update && ll == 1) {
        if (d.variable_probe && vp_side) {
          // the evolution dots need the updated V (second stream, in order) and, with variable
          // intensity, the updated probe (this stream): they follow on the second stream, beside
          // the probe-mode update, which touches neither evo, V nor probe mode 1
          evo_work = [&d, G, m, mm, ch, gch, gf, chi0, opg](hipStream_t st) -> hipError_t {
            const float2* page2 = d.variable_intensity ? d.probe : nullptr;
            hipLaunchKernelGGL(k_frame_dots, gch, dim3(NT), 0, st, opg, d.vmode, page2, chi0, m, d.part);
            hipLaunchKernelGGL(k_evo_update, dim3(1), dim3(1024), 0, st, d.part, ch, gf, G, mm, d.evo, d.n,
                               page2 ? 2 : 1, d.ftot);
            return hipGetLastError();
          };
        } else if (d.variable_probe) {
          {
            const int nb = (mm + NT - 1) / NT;
            double* p2 = d.part + (long)G * ch;  // |acc|^2 partials
            double* p3 = p2 + nb;                // |V|^2 partials
            LAUNCH(k_vp_resid, gch, dim3(NT), d.tmp, d.pu, d.pum, d.illum, d.roi, gf, m, d.cols, d.vmode, d.gscal,
                   d.part);
            LAUNCH(k_vp_coef, dim3(1), dim3(1024), d.part, ch, gf, G, mm, d.evo, d.n, d.ftot, d.gscal);
            LAUNCH(k_vp_acc, dim3(nb), dim3(NT), d.acc, d.tmp, G, mm, d.gscal, p2);
            LAUNCH(k_vp_apply, dim3(nb), dim3(NT), d.vmode, d.acc, mm, p2, nb, (double)G / d.n, p3);
            LAUNCH(k_vp_norm, dim3(nb), dim3(NT), d.vmode, mm, p3, nb);
          }
          // with variable intensity too, both evolution updates take one pass (neither changes
          // op, chi, the variable mode or the probe, and they write disjoint halves of evo)
          const float2* page2 = d.variable_intensity ? d.probe : nullptr;
          LAUNCH(k_frame_dots, gch, dim3(NT), opg, d.vmode, page2, chi0, m, d.part);
          LAUNCH(k_evo_update, dim3(1), dim3(1024), d.part, ch, gf, G, mm, d.evo, d.n, page2 ? 2 : 1, d.ftot);
          CHK();
        }
        if (d.variable_intensity && !d.variable_probe) {
          LAUNCH(k_frame_dots, gch, dim3(NT), opg, d.probe, nullptr, chi0, m, d.part);
          LAUNCH(k_evo_update, dim3(1), dim3(1024), d.part, ch, gf, G, mm, d.evo, d.n, 0, d.ftot);
          CHK();
        }
      }
    }
    // the evolution dots (inputs final since the step application; nothing runs between it and
    // here on this stream) and the next group's illumination probes fork at one point
    const bool pre_side = prefetch && g + 1 < n_groups;
    if ((evo_work || pre_side) && !point2_bound) TRY(fork_point());
    if (point2_bound && !evo_work && !pre_side) return hipErrorInvalidValue;  // a bound point left unused
    if (evo_work) TRY(fork(1, false, false, std::move(evo_work)));
    if (pre_side) {
      const int Gn = goffs[g + 2] - goffs[g + 1];
      TRY(fork(0, vp_on_side, false, [&d, m, Gn, gfn, var_any](hipStream_t st) -> hipError_t {
        TRY(fft_rows_pair(d.spec, d.probe, d.vmode, m, d.tw, st));
        return vp_from_row_spectra(d.vp, d.spec, d.evo, d.sub, gfn, d.n, var_any ? 1 : 0, Gn, m, d.tw, st);
      }));
    }
    MARK(MLS_T_MODES);
    if (defer_modes) {
      TRY(mls_mode_updates(d.probe, d.psi, opg, d.illum, d.roi, gf, K, G, m, d.cols, fw, d.gscal + MLS_S_BP,
                           d.pum_part, d.pum_fpc, d.row_waves, d.tw, s));
    } else if (fuse_modes) {
      LAUNCH(k_pum_modes, dim3((mm + NT - 1) / NT, K - 1), dim3(NT), d.probe, d.psi, opg, d.illum, d.roi, gf, G, m,
             d.cols, fw, d.gscal);
      CHK();
    }
    // the group's second-stream work in fork order (when not already enqueued by the second host
    // thread), then the joins before the next group's writes
    for (int i = 0; i < n_side; ++i) TRY(side_task(side_work[i])());
    MARK(MLS_T_JOIN);
    // one join on this stream: with both side streams in use, stream A first waits for B's end
    const bool u0 = side_used[0], u1 = side_used[1];
    side_used[0] = side_used[1] = false;
    n_side = n_points = 0;
    if (!u0 && !u1) return hipSuccess;
    const int l = u0 ? 0 : 1;
    auto record_join = [&d, u0, u1, l]() -> hipError_t {
      if (u0 && u1) {
        TRY(hipEventRecord(d.ev_join[1], d.side2));
        TRY(hipStreamWaitEvent(d.side, d.ev_join[1], 0));
      }
      return hipEventRecord(d.ev_join[l], l ? d.side2 : d.side);
    };
    if (defer_join) {
      if (d.enq) d.enq->post(record_join);
      else TRY(record_join());
      join_lane = l;
      join_G = G;
    } else {
      TRY(record_join());
      TRY(hipStreamWaitEvent(s, d.ev_join[l], 0));
    }
    return hipSuccess;
  };
  // Group graphs (not under the launch timer): a group's shape and the iteration's flags key
  // one captured graph, replayed by one launch per group.
  const bool graphs = d.use_graphs && !tm && d.side && d.gsel && d.goffs_dev;
  defer_join = !graphs;
  if (graphs) LAUNCH(k_set_int, dim3(1), dim3(1), d.gsel, 0);
  for (int g = 0; g < n_groups; ++g) {
    if (!graphs) {
      TRY(group_body(g, d.gframes + goffs[g], g + 1 < n_groups ? d.gframes + goffs[g + 1] : nullptr));
      continue;
    }
    const int G = goffs[g + 1] - goffs[g], Gn = g + 1 < n_groups ? goffs[g + 2] - goffs[g + 1] : 0;
    if (G >= 4096 || Gn >= 4096) return hipErrorInvalidValue;
    const uint64_t key = (uint64_t)G | (uint64_t)Gn << 12 | (uint64_t)(g > 0) << 24 |
                         (uint64_t)(iter >= d.pps) << 25 | (uint64_t)var_update << 26 | (uint64_t)probe_rec << 27 |
                         (uint64_t)obj_rec_iter << 28 | (uint64_t)(iter < 10) << 29 | (uint64_t)d.mod_cb << 30 |
                         (uint64_t)d.pum_fpc << 36 | (uint64_t)d.row_waves << 44 | (uint64_t)(g & 1) << 45;
    auto it = d.graphs.find(key);
    if (it == d.graphs.end()) {
      if (d.graphs.size() >= 64) mls_drop_graphs(d);
      hipGraph_t graph = nullptr;
      TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      hipError_t e = hipSuccess;
      hipLaunchKernelGGL(k_select_group, dim3(1), dim3(64), 0, s, d.gframes, d.goffs_dev, Gn > 0 ? 1 : 0, d.gmax,
                         d.gsel);
      e = hipGetLastError();
      if (e == hipSuccess) e = group_body(g, d.gsel + 1, d.gsel + 1 + d.gmax);
      const hipError_t e2 = hipStreamEndCapture(s, &graph);
      if (e == hipSuccess) e = e2;
      hipGraphExec_t exec = nullptr;
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      if (e != hipSuccess) return e;
      it = d.graphs.emplace(key, exec).first;
    }
    TRY(hipGraphLaunch(it->second, s));
  }
  if (join_lane >= 0) TRY(join_wait());  // the last group's
  MARK(MLS_T_ITER);
  if (K > d.n_rec) TRY(mls_ortho(d, s));
  MARK(-1);
  return hipSuccess;
}

hipError_t mls_iteration(MlsDev& d, int iter, const int* goffs, int n_groups, hipStream_t s, MlsTimer* tm) {
  const hipError_t e = mls_iteration_body(d, iter, goffs, n_groups, s, tm);
  // an error between a fork and its join leaves work on the side streams that still reads the
  // group buffers and writes the positions, V and the evolution: it finishes before the error is
  // reported, so that no later call overwrites what it reads
  if (e != hipSuccess) {
    if (d.enq) (void)d.enq->drain();  // nothing more is enqueued on the side streams
    if (d.side) (void)hipStreamSynchronize(d.side);
    if (d.side2) (void)hipStreamSynchronize(d.side2);
  }
  return e;
}

#undef MARK
#undef LAUNCH
#undef CHK
#undef TRY

}  // namespace ptyk
