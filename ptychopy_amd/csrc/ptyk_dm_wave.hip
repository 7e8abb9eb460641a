// DM iteration, wave-scheduled form for M in {128, 256, 512} (same equations as ptyk_dm.hip;
// replaces DM.cu:80-128 and the cuFFT callbacks of SampleKernels.cu:418-488).
//
// Column transforms are split as M = R0 * 64 (R0 = M/64):
//   pass A (row groups): X = RowFFT(2 P O_r - psi) for the R0 rows {n1 + 64 n2}, then the
//          radix-R0 step across those rows with twiddle W_M^(n1 k2); result row n1 + 64 k2.
//   pass B (64 x 64 blocks): the remaining 64-point column transform over rows 64 k2 .. 64 k2 + 63
//          (frequency row R0 k1 + k2), the modulus difference (A/|Phi| - 1) Phi / M^2, and the
//          inverse 64-point transform -- contiguous 512 B row segments, 32 KB of LDS.
//   pass C (row groups): inverse twiddle and radix-R0 step, inverse row FFTs: psi_new = Pf - x.
// Frames are processed in waves small enough that a wave's in-place intermediates stay in the
// on-package cache between passes.  Object and probe increments are accumulated per wave into
// FP64 buffers in frame order (deterministic), and applied once after the last wave, so every
// pass reads the object and probe as they were at the start of the iteration.
#include <hip/hip_ext.h>

#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

template <int M>
struct WaveCfg {
  static constexpr int L = 64;
  static constexpr int R0 = M / 64;
  static constexpr int RNB = plan_min_batch<M, PTYK_NT>();
  static constexpr int GPW = RNB / R0 > 0 ? RNB / R0 : 1;  // row groups per workgroup
  static constexpr int ROWS = GPW * R0;                     // rows held per workgroup
  static constexpr int CPB = M / 64;                        // 64-column chunks per row
  static constexpr int BPF = R0 * CPB;                      // 64 x 64 blocks per frame
};

// Global <-> LDS traffic of the row passes moves two adjacent complex values (16 B) per lane, in
// fully unrolled loops so that every load of a thread is in flight before the first LDS store.
template <int M>
struct RowIO {
  using C = WaveCfg<M>;
  static constexpr int PAIRS = C::ROWS * M / 2 / PTYK_NT;  // float4 per thread
  static_assert((C::ROWS * M / 2) % PTYK_NT == 0, "row tile must divide among the threads");
  // element e (even) of the workgroup's tile -> (frame, frame row, column); false past the end
  __device__ __forceinline__ static bool locate(int e, long f0, long n_groups, long* f, int* r, int* c) {
    const int lr = e / M;
    *c = e % M;
    const long G = (long)blockIdx.x * C::GPW + lr / C::R0;
    *f = f0 + G / C::L;
    *r = (int)(G % C::L) + C::L * (lr % C::R0);
    return G < n_groups;
  }
};

// The radix-R0 step across the rows {n1 + 64 n2} of each row group in LDS, with the twiddle
// W_M^(n1 k2) after the forward DFT or (conjugated) before the inverse one. For M < 512 the
// GPW M / NT steps of a thread are unrolled so that all their twiddle loads are in flight at
// once; at M = 512 (two steps) the runtime loop measured faster.
template <int M, bool INV>
__device__ __forceinline__ void cross_rows(float2* lds, const float2* __restrict__ tw, int tid) {
  using C = WaveCfg<M>;
  using L = SwzRowLayout<M, M>;
  constexpr int STEPS = C::GPW * M / PTYK_NT;
  static_assert(C::GPW * M % PTYK_NT == 0, "cross-row step must divide among the threads");
  auto step = [&](int t) __attribute__((always_inline)) {
    const int g = t / M, c = t % M;
    const int n1 = (int)(((long)blockIdx.x * C::GPW + g) % C::L);
    float2 w[C::R0];
#pragma unroll
    for (int k2 = 1; k2 < C::R0; ++k2) {
      w[k2] = tw[(n1 * k2 * (PTYK_TWN / M)) & (PTYK_TWN - 1)];
      if (INV) w[k2] = cconj(w[k2]);
    }
    float2 v[C::R0];
#pragma unroll
    for (int q = 0; q < C::R0; ++q) v[q] = lds[L::addr(g * C::R0 + q, c)];
    if (!INV) Dft<C::R0, false>::run(v);
#pragma unroll
    for (int k2 = 1; k2 < C::R0; ++k2) v[k2] = cmul(v[k2], w[k2]);
    if (INV) Dft<C::R0, true>::run(v);
#pragma unroll
    for (int q = 0; q < C::R0; ++q) lds[L::addr(g * C::R0 + q, c)] = v[q];
  };
  if constexpr (M < 512) {
#pragma unroll
    for (int it = 0; it < STEPS; ++it) step(tid + it * PTYK_NT);
  } else {
    for (int t = tid; t < C::GPW * M; t += PTYK_NT) step(t);
  }
}

// One wave's tile of a row pass (M <= 256): GPV whole row groups of one plane, ROWS = GPV R0
// rows {n1 + 64 q}, n1 = t GPV + g for tile t of the plane's T = 64 / GPV tiles.  Pair i of a
// lane is tile element 2 (lane + 64 i): local row lrow = g R0 + q, column lcol; a lane holds the
// same columns in every row of its row groups (pair (g R0 + q) HALVES + h), so the radix-R0 step
// across the rows runs in its registers with no exchange.
template <int M>
struct WaveTile {
  static constexpr int R0 = M / 64;
  static constexpr int GPV = plan_min_batch<M, 64>() / R0;
  static constexpr int ROWS = GPV * R0;
  static constexpr int PAIRS = ROWS * M / 128;  // float4 per lane
  static constexpr int HALVES = M / 128;        // column pairs per row of a lane
  static constexpr int T = 64 / GPV;
  static_assert(M == 128 || M == 256, "a wave tile of whole row groups: M = 128 or 256");
  static_assert(PAIRS == 8 && 64 % GPV == 0, "tile shape");
  __device__ __forceinline__ static int lrow(int lane, int i) { return (2 * (lane + 64 * i)) / M; }
  __device__ __forceinline__ static int lcol(int lane, int i) { return (2 * (lane + 64 * i)) % M; }
  __device__ __forceinline__ static int grow(int t, int lane, int i) {
    const int lr = lrow(lane, i);
    return t * GPV + lr / R0 + 64 * (lr % R0);
  }
  // cross_rows' arithmetic on the lane's pairs: forward DFT then W_M^(n1 k2), or the conjugated
  // twiddle then the inverse DFT
  template <bool INV>
  __device__ __forceinline__ static void cross(float4* v, int t, const float2* __restrict__ tw) {
#pragma unroll
    for (int g = 0; g < GPV; ++g) {
      const int n1 = t * GPV + g;
      float2 w[R0];
#pragma unroll
      for (int k2 = 1; k2 < R0; ++k2) {
        w[k2] = tw[(n1 * k2 * (PTYK_TWN / M)) & (PTYK_TWN - 1)];
        if (INV) w[k2] = cconj(w[k2]);
      }
#pragma unroll
      for (int h = 0; h < HALVES; ++h) {
        float2 a[R0], b[R0];
#pragma unroll
        for (int q = 0; q < R0; ++q) {
          const float4 x = v[(g * R0 + q) * HALVES + h];
          a[q] = make_float2(x.x, x.y);
          b[q] = make_float2(x.z, x.w);
        }
        if (!INV) {
          Dft<R0, false>::run(a);
          Dft<R0, false>::run(b);
        }
#pragma unroll
        for (int k2 = 1; k2 < R0; ++k2) {
          a[k2] = cmul(a[k2], w[k2]);
          b[k2] = cmul(b[k2], w[k2]);
        }
        if (INV) {
          Dft<R0, true>::run(a);
          Dft<R0, true>::run(b);
        }
#pragma unroll
        for (int q = 0; q < R0; ++q) v[(g * R0 + q) * HALVES + h] = make_float4(a[q].x, a[q].y, b[q].x, b[q].y);
      }
    }
  }
};

// Forward half of a row pass once its input rows are in LDS (the caller has not synchronised):
// row FFTs, the radix-R0 step across rows, and the store of the tile.
template <int M>
__device__ __forceinline__ void row_pass_forward(float2* lds, float2* __restrict__ psi, long f0, long n_groups,
                                                 const float2* __restrict__ tw, int tid) {
  using C = WaveCfg<M>;
  using IO = RowIO<M>;
  using L = SwzRowLayout<M, M>;
  __syncthreads();
  fft_lds<M, false, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
  cross_rows<M, false>(lds, tw, tid);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    long f;
    int r, c;
    if (IO::locate(e, f0, n_groups, &f, &r, &c))
      *reinterpret_cast<float4*>(psi + (f * M + r) * M + c) =
          *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
  }
}

// EXIT: the row input is DM's x = 2 P O_r - psi; otherwise psi itself (a plain forward row pass,
// used by the MLs modulus passes; obj, probe and roi are then not read).
template <int M, bool EXIT = true>
__global__ __launch_bounds__(PTYK_NT) void k_dmw_pass_a(float2* __restrict__ psi, const float2* __restrict__ obj,
                                                        int obj_cols, const float2* __restrict__ probe,
                                                        const int2* __restrict__ roi, long f0, long n_groups,
                                                        const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  using IO = RowIO<M>;
  using L = SwzRowLayout<M, M>;
  __shared__ float2 lds[C::ROWS * M];
  const int tid = threadIdx.x;
  // every row group of a workgroup belongs to one frame (L is a multiple of GPW), so its ROI is
  // one uniform load and the object addresses do not wait on a per-lane ROI load
  static_assert(C::L % C::GPW == 0, "a workgroup's row groups must share a frame");
  const int2 o = EXIT ? roi[f0 + ((long)blockIdx.x * C::GPW) / C::L] : make_int2(0, 0);
  // two halves of the tile, each with all of its loads in flight before its LDS stores
  constexpr int HALF = IO::PAIRS > 1 ? IO::PAIRS / 2 : 1;
#pragma unroll
  for (int h0 = 0; h0 < IO::PAIRS; h0 += HALF) {
    float4 sv[HALF], pv[HALF];
    float2 o0[HALF], o1[HALF];
#pragma unroll
    for (int u = 0; u < HALF; ++u) {
      const int e = 2 * (tid + (h0 + u) * PTYK_NT);
      long f;
      int r, c;
      if (IO::locate(e, f0, n_groups, &f, &r, &c)) {
        sv[u] = *reinterpret_cast<const float4*>(psi + (f * M + r) * M + c);
        if (EXIT) {
          const float2* orow = obj + (long)(o.x + r) * obj_cols + (o.y + c);
          pv[u] = *reinterpret_cast<const float4*>(probe + r * M + c);
          o0[u] = orow[0];
          o1[u] = orow[1];
        }
      } else {
        sv[u] = pv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        o0[u] = o1[u] = make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int u = 0; u < HALF; ++u) {
      const int e = 2 * (tid + (h0 + u) * PTYK_NT);
      if (EXIT) {
        const float2 a = cmul(make_float2(pv[u].x, pv[u].y), o0[u]);
        const float2 b = cmul(make_float2(pv[u].z, pv[u].w), o1[u]);
        *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) =
            make_float4(2.f * a.x - sv[u].x, 2.f * a.y - sv[u].y, 2.f * b.x - sv[u].z, 2.f * b.y - sv[u].w);
      } else {
        *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) = sv[u];
      }
    }
  }
  row_pass_forward<M>(lds, psi, f0, n_groups, tw, tid);
}

template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dmw_pass_b(float2* __restrict__ psi, const __half* __restrict__ amp,
                                                        const float* __restrict__ amp_scale, long f0,
                                                        const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  constexpr int PAIRS = 64 * 64 / 2 / PTYK_NT;
  __shared__ float2 lds[64 * 64];
  const int tid = threadIdx.x;
  const long f = f0 + blockIdx.x / C::BPF;
  const int rem = blockIdx.x % C::BPF;
  const int k2 = rem / C::CPB, c0 = (rem % C::CPB) * 64;
  float2* base = psi + (f * M + 64 * k2) * M + c0;
  {
    float4 v[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      v[i] = *reinterpret_cast<const float4*>(base + (e >> 6) * M + (e & 63));
    }
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) *reinterpret_cast<float4*>(&lds[2 * (tid + i * PTYK_NT)]) = v[i];
  }
  // amplitudes of this block's frequency rows, fetched before the transform so they are in
  // flight while it runs
  const float inv_m2 = 1.f / (float)(M * M);
  const float sc = amp_scale[f];
  const __half* a = amp + f * M * M + c0;
  __half2 av[PAIRS];
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    av[i] = *reinterpret_cast<const __half2*>(a + (long)(C::R0 * (e >> 6) + k2) * M + (e & 63));
  }
  __syncthreads();
  fft_lds<64, false, 64, PTYK_NT, ColLayout<64>, true>(lds, tw, tid);
  // each lane's pair of values moves as one 16-byte LDS access (two 8-byte accesses 16 bytes
  // apart across the lanes would meet bank conflicts)
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    const float2 amp2 = __half22float2(av[i]);
    const float4 q = *reinterpret_cast<const float4*>(&lds[e]);
    float2 d[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float A = (h ? amp2.y : amp2.x) * sc;
      const float2 ph = h ? make_float2(q.z, q.w) : make_float2(q.x, q.y);
      const float m2 = cabs2(ph);
      if (m2 > 0.f) {
        const float k = (A * rsqrtf(m2) - 1.f) * inv_m2;
        d[h] = make_float2(ph.x * k, ph.y * k);
      } else {
        d[h] = make_float2(A * inv_m2, 0.f);
      }
    }
    *reinterpret_cast<float4*>(&lds[e]) = make_float4(d[0].x, d[0].y, d[1].x, d[1].y);
  }
  __syncthreads();
  fft_lds<64, true, 64, PTYK_NT, ColLayout<64>, true>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    *reinterpret_cast<float4*>(base + (e >> 6) * M + (e & 63)) = *reinterpret_cast<const float4*>(&lds[e]);
  }
}

template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dmw_pass_c(float2* __restrict__ psi, long f0, long n_groups,
                                                        const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  using IO = RowIO<M>;
  using L = SwzRowLayout<M, M>;
  __shared__ float2 lds[C::ROWS * M];
  const int tid = threadIdx.x;
  {
    float4 v[IO::PAIRS];
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      long f;
      int r, c;
      v[i] = IO::locate(e, f0, n_groups, &f, &r, &c) ? *reinterpret_cast<const float4*>(psi + (f * M + r) * M + c)
                                                     : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) = v[i];
    }
  }
  __syncthreads();
  cross_rows<M, true>(lds, tw, tid);
  __syncthreads();
  fft_lds<M, true, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    long f;
    int r, c;
    if (IO::locate(e, f0, n_groups, &f, &r, &c))
      *reinterpret_cast<float4*>(psi + (f * M + r) * M + c) =
          *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
  }
}

#define PTYK_OBJ_CHUNK 256

// dO[p][c] += sum over this wave's frames covering object row p (ascending frame id) of the
// (optionally phase-clamped) increments s_O conj(P) psi_new.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dmw_obj_accum(double2* __restrict__ d_obj, int obj_cols, int row_lo,
                                                           const float2* __restrict__ psi,
                                                           const float2* __restrict__ probe,
                                                           const int2* __restrict__ roi,
                                                           const int* __restrict__ row_ptr,
                                                           const int* __restrict__ row_frames,
                                                           const float* __restrict__ factors, int phase_constraint) {
  const int p = row_lo + blockIdx.x;
  const int c = blockIdx.y * PTYK_OBJ_CHUNK + threadIdx.x;
  const float so = factors[1];
  double ax = 0.0, ay = 0.0;
  const int k0 = row_ptr[p], k1 = row_ptr[p + 1];
  // frames in ascending order, four at a time: all loads of a step are issued before the
  // (in-order) accumulation, so the sum is the same as one frame at a time
  constexpr int U = 4;
  for (int kb = k0; kb < k1; kb += U) {
    float2 pv[U], sv[U];
    bool act[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      act[u] = false;
      if (kb + u < k1) {
        const int j = row_frames[kb + u];
        const int2 r = roi[j];
        const int q = p - r.x, cc = c - r.y;
        act[u] = c < obj_cols && cc >= 0 && cc < M;
        if (act[u]) {
          pv[u] = probe[q * M + cc];
          sv[u] = psi[((long)j * M + q) * M + cc];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!act[u]) continue;
      float2 t = cscale(cmulcj(pv[u], sv[u]), so);
      if (phase_constraint) t = clamp_phase(t);
      ax += t.x;
      ay += t.y;
    }
  }
  if (c < obj_cols && k1 > k0) {
    double2 d = d_obj[(long)p * obj_cols + c];
    d.x += ax;
    d.y += ay;
    d_obj[(long)p * obj_cols + c] = d;
  }
}

// dP[q][c] += sum over frames [f0, f1) of conj(O[r_j + (q, c)]) psi_new_j[q][c].
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dmw_probe_accum(double2* __restrict__ d_probe,
                                                             const float2* __restrict__ psi,
                                                             const float2* __restrict__ obj, int obj_cols,
                                                             const int2* __restrict__ roi, long f0, long f1) {
  constexpr int CPT = M >= PTYK_NT ? M / PTYK_NT : 1;
  constexpr int RPW = M >= PTYK_NT ? 1 : PTYK_NT / M;
  const int q = blockIdx.x * RPW + (RPW > 1 ? threadIdx.x / M : 0);
  const int c0 = RPW > 1 ? threadIdx.x % M : threadIdx.x;
  double ax[CPT], ay[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) ax[k] = ay[k] = 0.0;
#pragma unroll 4
  for (long j = f0; j < f1; ++j) {
    const int2 r = roi[j];
    const float2* prow = psi + (j * M + q) * M;
    const float2* orow = obj + (long)(r.x + q) * obj_cols + r.y;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int c = c0 + k * PTYK_NT;
      const float2 t = cmulcj(orow[c], prow[c]);
      ax[k] += t.x;
      ay[k] += t.y;
    }
  }
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int c = c0 + k * PTYK_NT;
    double2 d = d_probe[q * M + c];
    d.x += ax[k];
    d.y += ay[k];
    d_probe[q * M + c] = d;
  }
}

// O += dO (one rounding), dO <- 0, per-block sum |O|^2.
__global__ __launch_bounds__(PTYK_NT) void k_dmw_obj_apply(float2* __restrict__ obj, double2* __restrict__ d_obj,
                                                           int obj_cols, double* __restrict__ block_sums) {
  __shared__ double red[PTYK_NT / 64];
  const int p = blockIdx.x, c = blockIdx.y * PTYK_OBJ_CHUNK + threadIdx.x;
  double s2 = 0.0;
  if (c < obj_cols) {
    const long i = (long)p * obj_cols + c;
    const double2 d = d_obj[i];
    float2 o = obj[i];
    o.x = (float)((double)o.x + d.x);
    o.y = (float)((double)o.y + d.y);
    obj[i] = o;
    d_obj[i] = make_double2(0.0, 0.0);
    s2 = (double)o.x * o.x + (double)o.y * o.y;
  }
  s2 = block_sum<PTYK_NT>(s2, red);
  if (threadIdx.x == 0) block_sums[(long)blockIdx.x * gridDim.y + blockIdx.y] = s2;
}

// P += s_P dP when updating, dP <- 0, per-block sum |P|^2.
__global__ __launch_bounds__(PTYK_NT) void k_dmw_probe_apply(float2* __restrict__ probe, double2* __restrict__ d_probe,
                                                             long mm, const float* __restrict__ factors, int update,
                                                             double* __restrict__ block_sums) {
  __shared__ double red[PTYK_NT / 64];
  const long i = (long)blockIdx.x * PTYK_NT + threadIdx.x;
  double s2 = 0.0;
  if (i < mm) {
    float2 p = probe[i];
    if (update) {
      const double2 d = d_probe[i];
      const double sp = (double)factors[0];
      p.x = (float)((double)p.x + sp * d.x);
      p.y = (float)((double)p.y + sp * d.y);
      probe[i] = p;
      d_probe[i] = make_double2(0.0, 0.0);
    }
    s2 = (double)p.x * p.x + (double)p.y * p.y;
  }
  s2 = block_sum<PTYK_NT>(s2, red);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = s2;
}

template <int M>
static hipError_t wave_impl(const DmWaveArgs& a, hipStream_t s, DmTimer* tm) {
  using C = WaveCfg<M>;
  const int cch = (a.obj_cols + PTYK_OBJ_CHUNK - 1) / PTYK_OBJ_CHUNK;
  constexpr int RPW = M >= PTYK_NT ? 1 : PTYK_NT / M;
  for (int w = 0; w < a.n_waves; ++w) {
    const long f0 = (long)w * a.wave, f1 = f0 + a.wave < a.n_frames ? f0 + a.wave : a.n_frames;
    const long groups = (f1 - f0) * C::L;
    const unsigned ga = (unsigned)((groups + C::GPW - 1) / C::GPW);
    const unsigned gb = (unsigned)((f1 - f0) * C::BPF);
    if (tm) tm->mark(s, DM_T_ROWS_FWD);
    hipLaunchKernelGGL((k_dmw_pass_a<M>), dim3(ga), dim3(PTYK_NT), 0, s, a.psi, a.obj, a.obj_cols, a.probe, a.roi, f0,
                       groups, a.tw);
    if (tm) tm->mark(s, DM_T_COLS);
    hipLaunchKernelGGL((k_dmw_pass_b<M>), dim3(gb), dim3(PTYK_NT), 0, s, a.psi, a.amp, a.amp_scale, f0, a.tw);
    if (tm) tm->mark(s, DM_T_ROWS_INV);
    hipLaunchKernelGGL((k_dmw_pass_c<M>), dim3(ga), dim3(PTYK_NT), 0, s, a.psi, f0, groups, a.tw);
    if (a.update_probe) {
      if (tm) tm->mark(s, DM_T_PROBE_GATHER);
      hipLaunchKernelGGL((k_dmw_probe_accum<M>), dim3(M / RPW), dim3(PTYK_NT), 0, s, a.d_probe, a.psi, a.obj,
                         a.obj_cols, a.roi, f0, f1);
    }
    const int lo = a.wave_row_lo[w], hi = a.wave_row_hi[w];
    if (hi > lo) {
      if (tm) tm->mark(s, DM_T_OBJ_GATHER);
      hipLaunchKernelGGL((k_dmw_obj_accum<M>), dim3(hi - lo, cch), dim3(PTYK_NT), 0, s, a.d_obj, a.obj_cols, lo,
                         a.psi, a.probe, a.roi, a.wave_row_ptr + (long)w * (a.obj_rows + 1), a.wave_row_frames,
                         a.factors, a.phase_constraint);
    }
  }
  if (tm) tm->mark(s, DM_T_PROBE_APPLY);
  const unsigned gpa = (unsigned)((M * M + PTYK_NT - 1) / PTYK_NT);
  hipLaunchKernelGGL(k_dmw_probe_apply, dim3(gpa), dim3(PTYK_NT), 0, s, a.probe, a.d_probe, (long)M * M, a.factors,
                     a.update_probe, a.probe_partial_sum);
  if (tm) tm->mark(s, DM_T_OBJ_APPLY);
  hipLaunchKernelGGL(k_dmw_obj_apply, dim3(a.obj_rows, cch), dim3(PTYK_NT), 0, s, a.obj, a.d_obj, a.obj_cols,
                     a.obj_partial);
  if (tm) tm->mark(s, DM_T_FINISH);
  hipLaunchKernelGGL((k_dm_finish<PTYK_NT>), dim3(1), dim3(PTYK_NT), 0, s, a.obj_partial, a.obj_rows * cch,
                     a.probe_partial_sum, (int)gpa, a.scal, a.factors);
  if (tm) tm->mark(s, -1);
  return hipGetLastError();
}

// ------------------------------------------------------------------ MLs modulus in three passes
// The MLs exit waves (K modes x G frames, mode-major planes) go through forward FFT, the
// multi-mode modulus step and the inverse FFT (MLS.cpp:119-140 with Diffractions.cpp:261-310) in
// three passes instead of five: the row passes above (plain input), and between them one pass
// over 64 x 16 blocks that holds every mode of a block, so the modulus, which needs
// sum_k |Psi_k|^2 at each pixel, runs between the forward and inverse 64-point column transforms.

// Frequency of FFT index i in cycles per sample, in [-0.5, 0.5) (the MLs spectral shift's).
__device__ __forceinline__ float shift_freq(int i, int m) {
  return (float)((i + (m >> 1)) & (m - 1)) / (float)m - 0.5f;
}

// Row pass of the MLs exit waves formed in place of their load (MLS.cpp:119-126 with
// d_readFromArray_c, utilitiesKernels.cu:1064-1084): plane k G + j, row r holds
// psi_k = P_k O_r, or O_r vp_j for mode 0 with the sub-pixel shift; mode 0's workgroups also
// write the object patch O_r (op, and op2 for the position step). Same arithmetic as the exit-wave
// kernel it replaces, so the transform sees the same values.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_mls_pass_a(float2* __restrict__ psi, MlsExitArgs x, int G,
                                                        long n_groups, const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  using IO = RowIO<M>;
  using L = SwzRowLayout<M, M>;
  static_assert(C::L % C::GPW == 0, "a workgroup's row groups must share a plane");
  __shared__ float2 lds[C::ROWS * M];
  const int tid = threadIdx.x;
  const long plane = ((long)blockIdx.x * C::GPW) / C::L;  // uniform: k G + j
  const int k = (int)(plane / G), j = (int)(plane % G);
  const int2 o = x.roi[x.gf[j]];
  const long mm = (long)M * M;
  const float2* pk = x.probe + (long)k * mm;
  const bool shifted = k == 0 && x.subpix;
  constexpr int HALF = IO::PAIRS > 1 ? IO::PAIRS / 2 : 1;
#pragma unroll
  for (int h0 = 0; h0 < IO::PAIRS; h0 += HALF) {
    float4 pv[HALF];
    float2 o0[HALF], o1[HALF];
#pragma unroll
    for (int u = 0; u < HALF; ++u) {
      const int e = 2 * (tid + (h0 + u) * PTYK_NT);
      long f;
      int r, c;
      if (IO::locate(e, 0, n_groups, &f, &r, &c)) {
        const float2* orow = x.obj + (long)(o.x + r) * x.cols + (o.y + c);
        o0[u] = orow[0];
        o1[u] = orow[1];
        pv[u] = shifted ? *reinterpret_cast<const float4*>(x.vp + (long)j * mm + r * M + c)
                        : *reinterpret_cast<const float4*>(pk + r * M + c);
      } else {
        pv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        o0[u] = o1[u] = make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int u = 0; u < HALF; ++u) {
      const int e = 2 * (tid + (h0 + u) * PTYK_NT);
      long f;
      int r, c;
      const bool in = IO::locate(e, 0, n_groups, &f, &r, &c);
      const float2 p0 = make_float2(pv[u].x, pv[u].y), p1 = make_float2(pv[u].z, pv[u].w);
      // k_psi's operand order: O * vp for the shifted mode 0, P * O otherwise
      const float2 a = shifted ? cmul(o0[u], p0) : cmul(p0, o0[u]);
      const float2 b = shifted ? cmul(o1[u], p1) : cmul(p1, o1[u]);
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) = make_float4(a.x, a.y, b.x, b.y);
      if (k == 0 && in) {
        const float4 ov = make_float4(o0[u].x, o0[u].y, o1[u].x, o1[u].y);
        *reinterpret_cast<float4*>(x.op + (long)j * mm + r * M + c) = ov;
        if (x.op2) *reinterpret_cast<float4*>(x.op2 + (long)j * mm + r * M + c) = ov;
      }
    }
  }
  row_pass_forward<M>(lds, psi, 0, n_groups, tw, tid);
  if (k == 0 && x.dx) {
    // the column derivative of the patch rows this workgroup wrote (read back by the threads
    // that wrote them): row transform, 2 pi i f_col / M, inverse row transform
    __syncthreads();  // the exit-wave stores have read the tile
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      long f;
      int r, c;
      IO::locate(e, 0, n_groups, &f, &r, &c);
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) =
          *reinterpret_cast<const float4*>(x.op + (long)j * mm + r * M + c);
    }
    __syncthreads();
    fft_lds<M, false, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
    const float inv = 1.f / (float)M;
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      const int c = e % M;
      float4* p = reinterpret_cast<float4*>(&lds[L::addr(e / M, c)]);
      const float4 t = *p;
      const float s0 = shift_freq(c, M) * inv, s1 = shift_freq(c + 1, M) * inv;
      const float2 a = cscale(make_float2(-t.y * 6.283185307179586f, t.x * 6.283185307179586f), s0);
      const float2 b = cscale(make_float2(-t.w * 6.283185307179586f, t.z * 6.283185307179586f), s1);
      *p = make_float4(a.x, a.y, b.x, b.y);
    }
    __syncthreads();
    fft_lds<M, true, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      long f;
      int r, c;
      IO::locate(e, 0, n_groups, &f, &r, &c);
      *reinterpret_cast<float4*>(x.dx + (long)j * mm + r * M + c) =
          *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
    }
  }
}

// The MLs sub-pixel shift x <- IFFT2(FFT2(x) exp(-2 pi i sign (sub.x f[col] + sub.y f[row]))) / M^2
// (Probe::shift_probe, Probe.cpp:1301-1331) is separable: the ramp is a product of a factor
// that depends on the column frequency only and one that depends on the row frequency only, so
// it is a shift of every row (transform, ramp, inverse transform along the row, / M) followed by
// a shift of every column.  Two passes over the planes instead of a 2D transform pair.

// exp(-2 pi i nom) t / M
__device__ __forceinline__ float2 shift_ramp(float2 t, float nom, float inv_m) {
  float sn, cs;
  sincosf(nom * -6.283185307179586f, &sn, &cs);
  return cscale(cmul(t, make_float2(cs, sn)), inv_m);
}

// Row shifts of the probe-update planes formed in place of their load: plane j, row r holds
// chi0_j conj(op_j) (the arithmetic of the probe-update kernel it replaces).  A workgroup's rows
// belong to one plane (NB divides M).
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_mls_pu_rows(float2* __restrict__ pu, const float2* __restrict__ chi,
                                                         const float2* __restrict__ op, const float2* __restrict__ sub,
                                                         const int* __restrict__ gf, float sign,
                                                         const float2* __restrict__ tw) {
  constexpr int NB = plan_min_batch<M, PTYK_NT>();
  constexpr int PAIRS = NB * M / 2 / PTYK_NT;
  static_assert(M % NB == 0 && (NB * M / 2) % PTYK_NT == 0, "row tile must divide the plane and the threads");
  using L = SwzRowLayout<M, M>;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * NB;
  const int j = (int)(row0 / M);
  {
    float4 av[PAIRS], bv[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const long g = row0 * M + 2 * (tid + i * PTYK_NT);
      av[i] = *reinterpret_cast<const float4*>(chi + g);
      bv[i] = *reinterpret_cast<const float4*>(op + g);
    }
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      const float4 a = av[i], b = bv[i];
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) =
          make_float4(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y, a.z * b.z + a.w * b.w, a.w * b.z - a.z * b.w);
    }
  }
  // a thread's columns are the same in every row it holds (2 NT is a multiple of M), so its
  // two ramp factors are computed once
  static_assert((2 * PTYK_NT) % M == 0, "a thread's columns must repeat across its rows");
  const float sx = sign * sub[gf[j]].x;
  const int c = (2 * tid) % M;
  float2 w0, w1;
  sincosf(sx * shift_freq(c, M) * -6.283185307179586f, &w0.y, &w0.x);
  sincosf(sx * shift_freq(c + 1, M) * -6.283185307179586f, &w1.y, &w1.x);
  __syncthreads();
  fft_lds<M, false, NB, PTYK_NT, L, false>(lds, tw, tid);
  const float inv_m = 1.f / (float)M;
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    float4* p = reinterpret_cast<float4*>(&lds[L::addr(e / M, c)]);
    const float4 t = *p;
    const float2 u0 = cscale(cmul(make_float2(t.x, t.y), w0), inv_m);
    const float2 u1 = cscale(cmul(make_float2(t.z, t.w), w1), inv_m);
    *p = make_float4(u0.x, u0.y, u1.x, u1.y);
  }
  __syncthreads();
  fft_lds<M, true, NB, PTYK_NT, L, false>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    *reinterpret_cast<float4*>(pu + row0 * M + e) = *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
  }
}

// The inverse row pass of mode 0 (k_dmw_pass_c's arithmetic, chi_0 written back for the rest of
// the group) followed, while the rows are still in LDS, by the row shifts of the probe update
// (k_mls_pu_rows's arithmetic on chi_0 conj(op)): one pass instead of two, and chi_0 is not read
// back.  Planes 0 .. G - 1 (mode 0) only; a workgroup's rows belong to one plane.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_mls_pass_c_pu(float2* __restrict__ psi, long n_groups,
                                                           float2* __restrict__ pu, const float2* __restrict__ op,
                                                           const float2* __restrict__ sub, const int* __restrict__ gf,
                                                           float sign, const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  using IO = RowIO<M>;
  using L = SwzRowLayout<M, M>;
  static_assert(C::L % C::GPW == 0, "a workgroup's row groups must share a plane");
  static_assert((2 * PTYK_NT) % M == 0, "a thread's columns must repeat across its rows");
  __shared__ float2 lds[C::ROWS * M];
  const int tid = threadIdx.x;
  const long mm = (long)M * M;
  const int j = (int)(((long)blockIdx.x * C::GPW) / C::L);  // plane = frame of the group
  {
    float4 v[IO::PAIRS];
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      long f;
      int r, c;
      v[i] = IO::locate(e, 0, n_groups, &f, &r, &c) ? *reinterpret_cast<const float4*>(psi + (f * M + r) * M + c)
                                                    : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) = v[i];
    }
  }
  // the object patch rows of the probe update, in flight during the inverse transform
  float4 bv[IO::PAIRS];
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    long f;
    int r, c;
    IO::locate(2 * (tid + i * PTYK_NT), 0, n_groups, &f, &r, &c);
    bv[i] = *reinterpret_cast<const float4*>(op + (long)j * mm + r * M + c);
  }
  __syncthreads();
  cross_rows<M, true>(lds, tw, tid);
  __syncthreads();
  fft_lds<M, true, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    long f;
    int r, c;
    IO::locate(e, 0, n_groups, &f, &r, &c);
    float4* p = reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]);
    const float4 a = *p, b = bv[i];
    *reinterpret_cast<float4*>(psi + (f * M + r) * M + c) = a;
    *p = make_float4(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y, a.z * b.z + a.w * b.w, a.w * b.z - a.z * b.w);
  }
  const float sx = sign * sub[gf[j]].x;
  const int c = (2 * tid) % M;
  float2 w0, w1;
  sincosf(sx * shift_freq(c, M) * -6.283185307179586f, &w0.y, &w0.x);
  sincosf(sx * shift_freq(c + 1, M) * -6.283185307179586f, &w1.y, &w1.x);
  __syncthreads();
  fft_lds<M, false, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
  const float inv_m = 1.f / (float)M;
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    float4* p = reinterpret_cast<float4*>(&lds[L::addr(e / M, c)]);
    const float4 t = *p;
    const float2 u0 = cscale(cmul(make_float2(t.x, t.y), w0), inv_m);
    const float2 u1 = cscale(cmul(make_float2(t.z, t.w), w1), inv_m);
    *p = make_float4(u0.x, u0.y, u1.x, u1.y);
  }
  __syncthreads();
  fft_lds<M, true, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    long f;
    int r, c2;
    IO::locate(e, 0, n_groups, &f, &r, &c2);
    *reinterpret_cast<float4*>(pu + (f * M + r) * M + c2) = *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
  }
}

// Wave-tile forms of the two MLs row passes (M <= 256): the same arithmetic per row (the same
// Stockham stages and the same radix-R0 step across the rows, so the results are bitwise those
// of k_mls_pass_a / k_mls_pass_c_pu), but each wave owns one tile of whole row groups
// (WaveTile) and exchanges through LDS with wave-local synchronisation only: the waves of a
// workgroup never wait for one another, so one wave's loads overlap another's transforms.
// Wave index blockIdx.x W + w covers planes x tiles; waves past the end return at once (no
// workgroup barrier follows).
template <int M, int W>
__global__ __launch_bounds__(64 * W) void k_mls_pass_a_w(float2* __restrict__ psi, MlsExitArgs x, int G, int p0,
                                                         int planes, const float2* __restrict__ tw) {
  using WT = WaveTile<M>;
  using L = SwzRowLayout<M, M>;
  constexpr int PAIRS = WT::PAIRS;
  __shared__ float2 lds[W][WT::ROWS * M];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wid = (int)blockIdx.x * W + w;
  if (wid >= planes * WT::T) return;
  const int plane = p0 + wid / WT::T, t = wid % WT::T;
  const int k = plane / G, j = plane % G;
  float2* tile = lds[w];
  const int2 o = x.roi[x.gf[j]];
  const long mm = (long)M * M;
  const float2* pk = x.probe + (long)k * mm;
  const bool shifted = k == 0 && x.subpix;
  // mode 0 in parts (the rest of the modes always whole): patches and dX only, or its exit waves
  // only from the patches already in op
  const bool exit_wave = !(k == 0 && x.mode0 == 1);
  const bool patches = k == 0 && x.mode0 != 2;
  const bool from_op = k == 0 && x.mode0 == 2;
  constexpr int HALF = PAIRS / 2;
#pragma unroll
  for (int h0 = 0; h0 < PAIRS; h0 += HALF) {
    float4 pv[HALF];
    float2 o0[HALF], o1[HALF];
#pragma unroll
    for (int u = 0; u < HALF; ++u) {
      const int r = WT::grow(t, lane, h0 + u), c = WT::lcol(lane, h0 + u);
      if (from_op) {
        const float4 ov = *reinterpret_cast<const float4*>(x.op + (long)j * mm + r * M + c);
        o0[u] = make_float2(ov.x, ov.y);
        o1[u] = make_float2(ov.z, ov.w);
      } else {
        const float2* orow = x.obj + (long)(o.x + r) * x.cols + (o.y + c);
        o0[u] = orow[0];
        o1[u] = orow[1];
      }
      if (exit_wave)
        pv[u] = shifted ? *reinterpret_cast<const float4*>(x.vp + (long)j * mm + r * M + c)
                        : *reinterpret_cast<const float4*>(pk + r * M + c);
    }
#pragma unroll
    for (int u = 0; u < HALF; ++u) {
      const int r = WT::grow(t, lane, h0 + u), c = WT::lcol(lane, h0 + u);
      if (exit_wave) {
        const float2 p0 = make_float2(pv[u].x, pv[u].y), p1 = make_float2(pv[u].z, pv[u].w);
        // k_psi's operand order: O * vp for the shifted mode 0, P * O otherwise
        const float2 a = shifted ? cmul(o0[u], p0) : cmul(p0, o0[u]);
        const float2 b = shifted ? cmul(o1[u], p1) : cmul(p1, o1[u]);
        *reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, h0 + u), c)]) = make_float4(a.x, a.y, b.x, b.y);
      }
      if (patches) {
        const float4 ov = make_float4(o0[u].x, o0[u].y, o1[u].x, o1[u].y);
        *reinterpret_cast<float4*>(x.op + (long)j * mm + r * M + c) = ov;
        if (x.op2) *reinterpret_cast<float4*>(x.op2 + (long)j * mm + r * M + c) = ov;
      }
    }
  }
  if (exit_wave) {
    lds_sync<true>();
    fft_lds<M, false, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
    {
      float4 v[PAIRS];
#pragma unroll
      for (int i = 0; i < PAIRS; ++i)
        v[i] = *reinterpret_cast<const float4*>(&tile[L::addr(WT::lrow(lane, i), WT::lcol(lane, i))]);
      WT::template cross<false>(v, t, tw);
      float2* dst = psi + (long)plane * mm;
#pragma unroll
      for (int i = 0; i < PAIRS; ++i)
        *reinterpret_cast<float4*>(dst + (long)WT::grow(t, lane, i) * M + WT::lcol(lane, i)) = v[i];
    }
  }
  if (patches && x.dx) {
    // the column derivative of the patch rows this wave wrote (read back by the lanes that wrote
    // them): row transform, 2 pi i f_col / M, inverse row transform
    lds_sync<true>();
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int c = WT::lcol(lane, i);
      *reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, i), c)]) =
          *reinterpret_cast<const float4*>(x.op + (long)j * mm + (long)WT::grow(t, lane, i) * M + c);
    }
    lds_sync<true>();
    fft_lds<M, false, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
    const float inv = 1.f / (float)M;
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int c = WT::lcol(lane, i);
      float4* p = reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, i), c)]);
      const float4 v = *p;
      const float s0 = shift_freq(c, M) * inv, s1 = shift_freq(c + 1, M) * inv;
      const float2 a = cscale(make_float2(-v.y * 6.283185307179586f, v.x * 6.283185307179586f), s0);
      const float2 b = cscale(make_float2(-v.w * 6.283185307179586f, v.z * 6.283185307179586f), s1);
      *p = make_float4(a.x, a.y, b.x, b.y);
    }
    lds_sync<true>();
    fft_lds<M, true, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int c = WT::lcol(lane, i);
      *reinterpret_cast<float4*>(x.dx + (long)j * mm + (long)WT::grow(t, lane, i) * M + c) =
          *reinterpret_cast<const float4*>(&tile[L::addr(WT::lrow(lane, i), c)]);
    }
  }
}

template <int M, int W>
__global__ __launch_bounds__(64 * W) void k_mls_pass_c_pu_w(float2* __restrict__ psi, int G, float2* __restrict__ pu,
                                                            const float2* __restrict__ op,
                                                            const float2* __restrict__ sub, const int* __restrict__ gf,
                                                            float sign, const float2* __restrict__ tw) {
  using WT = WaveTile<M>;
  using L = SwzRowLayout<M, M>;
  constexpr int PAIRS = WT::PAIRS;
  __shared__ float2 lds[W][WT::ROWS * M];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wid = (int)blockIdx.x * W + w;
  if (wid >= G * WT::T) return;
  const int j = wid / WT::T, t = wid % WT::T;
  float2* tile = lds[w];
  const long mm = (long)M * M;
  float2* pl = psi + (long)j * mm;
  const float2* ol = op + (long)j * mm;
  {
    float4 v[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
      v[i] = *reinterpret_cast<const float4*>(pl + (long)WT::grow(t, lane, i) * M + WT::lcol(lane, i));
    WT::template cross<true>(v, t, tw);
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
      *reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, i), WT::lcol(lane, i))]) = v[i];
  }
  // the object patch rows of the probe update, in flight during the inverse transform
  float4 bv[PAIRS];
#pragma unroll
  for (int i = 0; i < PAIRS; ++i)
    bv[i] = *reinterpret_cast<const float4*>(ol + (long)WT::grow(t, lane, i) * M + WT::lcol(lane, i));
  lds_sync<true>();
  fft_lds<M, true, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int c = WT::lcol(lane, i);
    float4* p = reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, i), c)]);
    const float4 a = *p, b = bv[i];
    *reinterpret_cast<float4*>(pl + (long)WT::grow(t, lane, i) * M + c) = a;
    *p = make_float4(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y, a.z * b.z + a.w * b.w, a.w * b.z - a.z * b.w);
  }
  const float sx = sign * sub[gf[j]].x;
  float2 w0[WT::HALVES], w1[WT::HALVES];
#pragma unroll
  for (int h = 0; h < WT::HALVES; ++h) {
    const int c = WT::lcol(lane, h);
    sincosf(sx * shift_freq(c, M) * -6.283185307179586f, &w0[h].y, &w0[h].x);
    sincosf(sx * shift_freq(c + 1, M) * -6.283185307179586f, &w1[h].y, &w1[h].x);
  }
  lds_sync<true>();
  fft_lds<M, false, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
  const float inv_m = 1.f / (float)M;
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int h = i % WT::HALVES;
    float4* p = reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, i), WT::lcol(lane, i))]);
    const float4 v = *p;
    const float2 u0 = cscale(cmul(make_float2(v.x, v.y), w0[h]), inv_m);
    const float2 u1 = cscale(cmul(make_float2(v.z, v.w), w1[h]), inv_m);
    *p = make_float4(u0.x, u0.y, u1.x, u1.y);
  }
  lds_sync<true>();
  fft_lds<M, true, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
  float2* dst = pu + (long)j * mm;
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int c = WT::lcol(lane, i);
    *reinterpret_cast<float4*>(dst + (long)WT::grow(t, lane, i) * M + c) =
        *reinterpret_cast<const float4*>(&tile[L::addr(WT::lrow(lane, i), c)]);
  }
}
constexpr int ROW_WAVES = 4;

// Column shifts in place: a workgroup takes NC columns (whole 16-column strips) of plane j.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_mls_shift_cols(float2* __restrict__ x, const float2* __restrict__ sub,
                                                            const int* __restrict__ gf, float sign,
                                                            const float2* __restrict__ tw) {
  constexpr int NC = (plan_min_batch<M, PTYK_NT>() + 15) / 16 * 16;
  constexpr int PAIRS = NC * M / 2 / PTYK_NT;
  static_assert((NC * M / 2) % PTYK_NT == 0, "column tile must divide among the threads");
  __shared__ float2 lds[NC * M];
  const int tid = threadIdx.x;
  const int j = blockIdx.y;
  float2* base = x + (long)j * M * M + blockIdx.x * NC;
  auto off = [&](int e) {  // e: strip-major element index of the tile (even)
    return (long)((e >> 4) % M) * M + (e / (M * 16)) * 16 + (e & 15);
  };
  {
    float4 v[PAIRS];
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) v[k] = *reinterpret_cast<const float4*>(base + off(2 * (tid + k * PTYK_NT)));
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) *reinterpret_cast<float4*>(&lds[2 * (tid + k * PTYK_NT)]) = v[k];
  }
  const float sy = sign * sub[gf[j]].y;
  __syncthreads();
  fft_lds<M, false, NC, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
  const float inv_m = 1.f / (float)M;
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const int e = 2 * (tid + k * PTYK_NT);
    const float nom = sy * shift_freq((e >> 4) % M, M);  // both values of the pair share the row
    float4* p = reinterpret_cast<float4*>(&lds[e]);
    const float4 t = *p;
    const float2 u0 = shift_ramp(make_float2(t.x, t.y), nom, inv_m);
    const float2 u1 = shift_ramp(make_float2(t.z, t.w), nom, inv_m);
    *p = make_float4(u0.x, u0.y, u1.x, u1.y);
  }
  __syncthreads();
  fft_lds<M, true, NC, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const int e = 2 * (tid + k * PTYK_NT);
    *reinterpret_cast<float4*>(base + off(e)) = *reinterpret_cast<const float4*>(&lds[e]);
  }
}

// K modes of a 64 x CB block as one batch: sequence b = mode * CB + column, element e = row.
template <int CB>
struct ModeColLayout {
  __device__ __forceinline__ static int addr(int b, int e) { return (b / CB) * (64 * CB) + e * CB + (b % CB); }
};

// Per block: Psi_k (64-point columns of rows 64 k2 .. 64 k2 + 63 after the row pass, i.e. the
// frequency rows R0 k1 + k2), s = |sum_k |Psi_k|^2|^(1/2), error partial 4 (A - s)^2, Psi_k *=
// (A / (s + 1e-9) - 1) / M^2 (k_modulus's arithmetic), inverse columns. One wave per mode (64 K
// threads) over 64 x 16 blocks: 16 KiB of LDS per mode, so four workgroups share a CU at K = 5.
// One error partial per block: part[j * BPF + block].  CB: the block's column width.
template <int M, int K, int CB>
__global__ __launch_bounds__(64 * K) void k_mls_pass_b(float2* __restrict__ psi, int G, const __half* __restrict__ amp,
                                                       const float* __restrict__ amp_scale, const int* __restrict__ gf,
                                                       double* __restrict__ part, const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  constexpr int BT = 64 * K;
  constexpr int NCB = M / CB, BPF = C::R0 * NCB;
  constexpr int BLK = 64 * CB;         // elements per mode
  constexpr int PAIRS = BLK / 2 / 64;  // float4 per lane of the mode's wave
  extern __shared__ float2 lds[];      // K x 64 x CB (mode, row, column)
  __shared__ double red[K];
  const int tid = threadIdx.x;
  const int wk = tid >> 6, ln = tid & 63;  // this wave's mode, lane
  const int j = blockIdx.x / BPF, blk = blockIdx.x % BPF;
  const int k2 = blk / NCB, c0 = (blk % NCB) * CB;
  const long mm = (long)M * M;
  {
    const float2* base = psi + ((long)wk * G + j) * mm + (long)(64 * k2) * M + c0;
    float4 v[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (ln + i * 64);
      v[i] = *reinterpret_cast<const float4*>(base + (long)(e / CB) * M + (e % CB));
    }
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) *reinterpret_cast<float4*>(&lds[wk * BLK + 2 * (ln + i * 64)]) = v[i];
  }
  // amplitudes of this thread's pixels (frequency row R0 k1 + k2), in flight during the transform
  constexpr int PIX = (BLK + BT - 1) / BT;
  const int f = gf[j];
  const float sc = amp_scale[f];
  float av[PIX];
#pragma unroll
  for (int p = 0; p < PIX; ++p) {
    const int idx = tid + p * BT;
    av[p] = idx < BLK ? __half2float(amp[(long)f * mm + (long)(C::R0 * (idx / CB) + k2) * M + c0 + idx % CB]) * sc
                      : 0.f;
  }
  // each wave transforms its own mode's columns with wave-local exchanges; the modulus, which
  // reads every mode at a pixel, waits for all of them
  lds_sync<true>();
  fft_lds<64, false, CB, 64, ModeColLayout<CB>, true, true, true>(lds + wk * BLK, tw, ln);
  __syncthreads();
  const float inv_mm = 1.f / (float)mm;
  double err = 0.0;
#pragma unroll
  for (int p = 0; p < PIX; ++p) {
    const int idx = tid + p * BT;
    if (idx >= BLK) break;
    float s2 = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k) s2 += cabs2(lds[k * BLK + idx]);
    const float ap = sqrtf(s2);
    const double d = (double)(av[p] - ap);
    err += d * d * 4.0;
    const float R = (float)((double)av[p] / ((double)ap + 1e-9) - 1.0) * inv_mm;
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * BLK + idx] = cscale(lds[k * BLK + idx], R);
  }
  __syncthreads();
  fft_lds<64, true, CB, 64, ModeColLayout<CB>, true, true, true>(lds + wk * BLK, tw, ln);
  {
    float2* base = psi + ((long)wk * G + j) * mm + (long)(64 * k2) * M + c0;
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (ln + i * 64);
      *reinterpret_cast<float4*>(base + (long)(e / CB) * M + (e % CB)) =
          *reinterpret_cast<const float4*>(&lds[wk * BLK + e]);
    }
  }
  err = block_sum<BT>(err, red);
  if (tid == 0) part[(long)j * BPF + blk] = err;
}

// A launch timed by its own dispatch's start/stop events when the timer has a pair free (timed
// iterations), a plain launch otherwise.
template <typename F, typename... A>
static void launch_own(LaunchTimer* tm, int category, F kernel, dim3 grid, dim3 block, size_t shm, hipStream_t s,
                       A... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (tm && tm->own(category, &e0, &e1))
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)shm, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, shm, s, args...);
}

template <int M, int K, int CB>
static hipError_t mls_modulus_k(float2* psi, int G, bool row_waves, const __half* amp, const float* amp_scale,
                                const int* gf, double* part, const float2* tw, const MlsExitArgs* exit, int c_modes,
                                const MlsShiftArgs* shift, hipStream_t s, MlsTimer* tm) {
  using C = WaveCfg<M>;
  constexpr int BPF = C::R0 * (M / CB);
  const long groups = (long)K * G * C::L, groups_c = (long)c_modes * G * C::L;
  const unsigned ga = (unsigned)((groups + C::GPW - 1) / C::GPW);
  const unsigned gc = (unsigned)((groups_c + C::GPW - 1) / C::GPW);
  const size_t lds = (size_t)K * 64 * CB * sizeof(float2);
  static bool lds_set = false;  // the attribute is per kernel and process: set it once
  if (!lds_set) {
    hipError_t e = hipFuncSetAttribute((const void*)k_mls_pass_b<M, K, CB>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    lds_set = true;
  }
  // wave-tile row passes (M <= 256): one wave per tile of whole row groups
  constexpr bool WAVE_OK = M <= 256;
  const bool waves = WAVE_OK && row_waves;
  const int wt = WaveTile<WAVE_OK ? M : 256>::T;
  if (exit && (exit->p0 != 0 || exit->np != 0 || exit->mode0 != 0) && !waves) return hipErrorInvalidValue;
  if (tm) tm->mark(s, MLS_T_PASS_A);
  if (exit && waves) {
    const int np = exit->np ? exit->np : K * G;
    if constexpr (WAVE_OK)
      launch_own(tm, MLS_T_OWN_A, k_mls_pass_a_w<M, ROW_WAVES>,
                 dim3((unsigned)((np * wt + ROW_WAVES - 1) / ROW_WAVES)), dim3(64 * ROW_WAVES), 0, s, psi, *exit, G,
                 exit->p0, np, tw);
  } else if (exit)
    launch_own(tm, MLS_T_OWN_A, k_mls_pass_a<M>, dim3(ga), dim3(PTYK_NT), 0, s, psi, *exit, G, groups, tw);
  else
    hipLaunchKernelGGL((k_dmw_pass_a<M, false>), dim3(ga), dim3(PTYK_NT), 0, s, psi, nullptr, 0, nullptr, nullptr, 0L,
                       groups, tw);
  if (tm) tm->mark(s, MLS_T_PASS_B);
  launch_own(tm, MLS_T_OWN_B, k_mls_pass_b<M, K, CB>, dim3((unsigned)(G * BPF)), dim3(64 * K), lds, s, psi, G, amp,
             amp_scale, gf, part, tw);
  if (tm) tm->mark(s, MLS_T_PASS_C);
  if (shift && c_modes == 1 && waves) {
    if constexpr (WAVE_OK)
      launch_own(tm, MLS_T_OWN_C, k_mls_pass_c_pu_w<M, ROW_WAVES>,
                 dim3((unsigned)((G * wt + ROW_WAVES - 1) / ROW_WAVES)), dim3(64 * ROW_WAVES), 0, s, psi, G,
                 shift->pu, shift->op, shift->sub, gf, shift->sign, tw);
  } else if (shift && c_modes == 1)
    launch_own(tm, MLS_T_OWN_C, k_mls_pass_c_pu<M>, dim3(gc), dim3(PTYK_NT), 0, s, psi, groups_c, shift->pu,
               shift->op, shift->sub, gf, shift->sign, tw);
  else
    launch_own(tm, MLS_T_OWN_C, k_dmw_pass_c<M>, dim3(gc), dim3(PTYK_NT), 0, s, psi, 0L, groups_c, tw);
  return hipGetLastError();
}

template <int M, int K>
static hipError_t mls_modulus_kc(float2* psi, int G, int cb, bool rw, const __half* amp, const float* amp_scale,
                                 const int* gf, double* part, const float2* tw, const MlsExitArgs* exit, int cm,
                                 const MlsShiftArgs* shift, hipStream_t s, MlsTimer* tm) {
  switch (cb) {
    case 8: return mls_modulus_k<M, K, 8>(psi, G, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 16: return mls_modulus_k<M, K, 16>(psi, G, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 32: return mls_modulus_k<M, K, 32>(psi, G, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    default: return hipErrorInvalidValue;
  }
}

template <int M>
static hipError_t mls_modulus_m(float2* psi, int K, int G, int cb, bool rw, const __half* amp, const float* amp_scale,
                                const int* gf, double* part, const float2* tw, const MlsExitArgs* exit, int cm,
                                const MlsShiftArgs* shift, hipStream_t s, MlsTimer* tm) {
  switch (K) {
    case 1: return mls_modulus_kc<M, 1>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 2: return mls_modulus_kc<M, 2>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 3: return mls_modulus_kc<M, 3>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 4: return mls_modulus_kc<M, 4>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 5: return mls_modulus_kc<M, 5>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 6: return mls_modulus_kc<M, 6>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 7: return mls_modulus_kc<M, 7>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    case 8: return mls_modulus_kc<M, 8>(psi, G, cb, rw, amp, amp_scale, gf, part, tw, exit, cm, shift, s, tm);
    default: return hipErrorInvalidValue;
  }
}

hipError_t mls_exit_rows(float2* psi, int G, int m, const MlsExitArgs* exit, int p0, int np, const float2* tw,
                         hipStream_t s, LaunchTimer* tm) {
  if (!exit || p0 < 0 || np < 1) return hipErrorInvalidValue;
  auto go = [&](auto kern, int T) {
    launch_own(tm, MLS_T_OWN_A, kern, dim3((unsigned)((np * T + ROW_WAVES - 1) / ROW_WAVES)), dim3(64 * ROW_WAVES), 0,
               s, psi, *exit, G, p0, np, tw);
    return hipGetLastError();
  };
  switch (m) {
    case 128: return go(k_mls_pass_a_w<128, ROW_WAVES>, WaveTile<128>::T);
    case 256: return go(k_mls_pass_a_w<256, ROW_WAVES>, WaveTile<256>::T);
    default: return hipErrorInvalidValue;
  }
}

int mls_modulus_blocks(int m, int k, int cb) {
  if ((m != 128 && m != 256 && m != 512) || k < 1 || k > 8 || (cb != 8 && cb != 16 && cb != 32)) return 0;
  return (m / 64) * (m / cb);
}

hipError_t mls_modulus_passes(float2* psi, int K, int G, int m, int cb, bool row_waves, const __half* amp,
                              const float* amp_scale,
                              const int* gf, double* part, const float2* tw, const MlsExitArgs* exit, int c_modes,
                              const MlsShiftArgs* shift, hipStream_t s, MlsTimer* tm) {
  if (c_modes < 1 || c_modes > K) return hipErrorInvalidValue;
  switch (mls_modulus_blocks(m, K, cb) ? m : 0) {
    case 128:
      return mls_modulus_m<128>(psi, K, G, cb, row_waves, amp, amp_scale, gf, part, tw, exit, c_modes, shift, s, tm);
    case 256:
      return mls_modulus_m<256>(psi, K, G, cb, row_waves, amp, amp_scale, gf, part, tw, exit, c_modes, shift, s, tm);
    case 512:
      return mls_modulus_m<512>(psi, K, G, cb, row_waves, amp, amp_scale, gf, part, tw, exit, c_modes, shift, s, tm);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------------------ MLs probe modes 2..K
// Workgroup (row tile t, mode k, chunk z): the inverse row pass of k_dmw_pass_c for the tile of
// each frame j of the chunk in turn (plane k G + j), then at every pixel of the tile
// acc += w_j chi conj(op_j), w_j = p / (p + fw) with p the illumination at the frame's ROI (the
// arithmetic of k_pum_modes); the chunk's sums go to part[(z (K - 1) + k - 1) M^2 + pixel].
// The chi_k planes are never written back: nothing else reads them.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_pum_rows(const float2* __restrict__ psi, const float2* __restrict__ op,
                                                      const float* __restrict__ illum, const int2* __restrict__ roi,
                                                      const int* __restrict__ gf, int K, int G, int cols, float fw,
                                                      float2* __restrict__ part, int fpc,
                                                      const float2* __restrict__ tw) {
  using C = WaveCfg<M>;
  using IO = RowIO<M>;
  using L = SwzRowLayout<M, M>;
  __shared__ float2 lds[C::ROWS * M];
  __shared__ float4 acc[IO::PAIRS * PTYK_NT];  // per-thread sums (LDS keeps them out of the registers)
  const int tid = threadIdx.x;
  const int k = blockIdx.y + 1;
  const int j0 = blockIdx.z * fpc, j1 = min(G, j0 + fpc);
  const long mm = (long)M * M;
  constexpr long NOLIM = 1L << 40;  // locate's bound: the tile always lies inside its plane
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) acc[i * PTYK_NT + tid] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int j = j0; j < j1; ++j) {
    const long plane = (long)k * G + j;
    if (j > j0) __syncthreads();  // the previous frame's products have read the tile
    float4 v[IO::PAIRS];
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      long f;
      int r, c;
      IO::locate(2 * (tid + i * PTYK_NT), plane, NOLIM, &f, &r, &c);
      v[i] = *reinterpret_cast<const float4*>(psi + (f * M + r) * M + c);
    }
    const int2 o = roi[gf[j]];
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      *reinterpret_cast<float4*>(&lds[L::addr(e / M, e % M)]) = v[i];
    }
    __syncthreads();
    cross_rows<M, true>(lds, tw, tid);
    __syncthreads();
    fft_lds<M, true, C::ROWS, PTYK_NT, L, false>(lds, tw, tid);
#pragma unroll
    for (int i = 0; i < IO::PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      long f;
      int r, c;
      IO::locate(e, plane, NOLIM, &f, &r, &c);
      const float4 bq = *reinterpret_cast<const float4*>(op + (long)j * mm + r * M + c);
      const float* ip = illum + (long)(o.x + r) * cols + o.y + c;
      const float p0 = ip[0], p1 = ip[1];
      const float4 u = *reinterpret_cast<const float4*>(&lds[L::addr(e / M, e % M)]);
      const float w0 = p0 / (p0 + fw), w1 = p1 / (p1 + fw);
      const float2 d0 = make_float2(u.x * bq.x + u.y * bq.y, u.y * bq.x - u.x * bq.y);
      const float2 d1 = make_float2(u.z * bq.z + u.w * bq.w, u.w * bq.z - u.z * bq.w);
      float4 t = acc[i * PTYK_NT + tid];
      t.x += w0 * d0.x;
      t.y += w0 * d0.y;
      t.z += w1 * d1.x;
      t.w += w1 * d1.y;
      acc[i * PTYK_NT + tid] = t;
    }
  }
  float2* out = part + ((long)blockIdx.z * (K - 1) + (k - 1)) * mm;
#pragma unroll
  for (int i = 0; i < IO::PAIRS; ++i) {
    long f;
    int r, c;
    IO::locate(2 * (tid + i * PTYK_NT), 0, NOLIM, &f, &r, &c);
    *reinterpret_cast<float4*>(out + r * M + c) = acc[i * PTYK_NT + tid];
  }
}

// The same sums with one wave per frame tile (M <= 256): workgroup (tile t, mode k, chunk z) of
// W waves; wave w takes the frames j0 + w, j0 + w + W, ... of the chunk, each through its own
// LDS tile of GPV row groups (ROWS = GPV R0 rows {n1 + 64 q}, n1 = t GPV + g): the radix-R0 step
// across the rows in registers (each lane loads whole columns of its row groups), the inverse
// row FFT with wave-local exchanges, then the weighted products into per-lane sums.  No
// workgroup barrier runs until the chunk is done, so the waves' loads and transforms overlap
// one another; the W per-wave sums then add in wave order (a fixed order: deterministic).
template <int M, int W>
__global__ __launch_bounds__(64 * W) void k_pum_waves(const float2* __restrict__ psi, const float2* __restrict__ op,
                                                      const float* __restrict__ illum, const int2* __restrict__ roi,
                                                      const int* __restrict__ gf, int K, int G, int cols, float fw,
                                                      float2* __restrict__ part, int fpc,
                                                      const float2* __restrict__ tw) {
  using WT = WaveTile<M>;
  using L = SwzRowLayout<M, M>;
  constexpr int PAIRS = WT::PAIRS;
  __shared__ float2 lds[W][WT::ROWS * M];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x, k = blockIdx.y + 1;
  const int j0 = blockIdx.z * fpc, j1 = min(G, j0 + fpc);
  const long mm = (long)M * M;
  float2* tile = lds[w];
  float4 acc[PAIRS];
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int j = j0 + w; j < j1; j += W) {
    const float2* src = psi + ((long)k * G + j) * mm;
    float4 v[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
      v[i] = *reinterpret_cast<const float4*>(src + (long)WT::grow(t, lane, i) * M + WT::lcol(lane, i));
    WT::template cross<true>(v, t, tw);
    lds_sync<true>();  // the previous frame's products have read the tile
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
      *reinterpret_cast<float4*>(&tile[L::addr(WT::lrow(lane, i), WT::lcol(lane, i))]) = v[i];
    lds_sync<true>();
    fft_lds<M, true, WT::ROWS, 64, L, false, true, true>(tile, tw, lane);
    const int2 o = roi[gf[j]];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int r = WT::grow(t, lane, i), c = WT::lcol(lane, i);
      const float4 bq = *reinterpret_cast<const float4*>(op + (long)j * mm + (long)r * M + c);
      const float* ip = illum + (long)(o.x + r) * cols + o.y + c;
      const float p0 = ip[0], p1 = ip[1];
      const float4 u = *reinterpret_cast<const float4*>(&tile[L::addr(WT::lrow(lane, i), c)]);
      const float w0 = p0 / (p0 + fw), w1 = p1 / (p1 + fw);
      const float2 d0 = make_float2(u.x * bq.x + u.y * bq.y, u.y * bq.x - u.x * bq.y);
      const float2 d1 = make_float2(u.z * bq.z + u.w * bq.w, u.w * bq.z - u.z * bq.w);
      acc[i].x += w0 * d0.x;
      acc[i].y += w0 * d0.y;
      acc[i].z += w1 * d1.x;
      acc[i].w += w1 * d1.y;
    }
  }
  // the waves' sums in wave order
  __syncthreads();
  float4* slots = reinterpret_cast<float4*>(tile);  // slot lane + 64 i of this wave's tile
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) slots[lane + 64 * i] = acc[i];
  __syncthreads();
  float2* out = part + ((long)blockIdx.z * (K - 1) + (k - 1)) * mm;
  for (int sl = threadIdx.x; sl < 64 * PAIRS; sl += 64 * W) {
    float4 a = reinterpret_cast<const float4*>(lds[0])[sl];
#pragma unroll
    for (int u = 1; u < W; ++u) {
      const float4 x = reinterpret_cast<const float4*>(lds[u])[sl];
      a.x += x.x;
      a.y += x.y;
      a.z += x.z;
      a.w += x.w;
    }
    const int i = sl >> 6, ln = sl & 63;
    *reinterpret_cast<float4*>(out + (long)WT::grow(t, ln, i) * M + WT::lcol(ln, i)) = a;
  }
}
constexpr int PUM_WAVES = 4;

// P_k += (sum over chunks in order) / G * beta for modes 2..K.
__global__ void k_pum_finish(float2* __restrict__ probe, const float2* __restrict__ part, int K, int chunks, int G,
                             long mm, const double* __restrict__ beta) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const int k = blockIdx.y + 1;
  float2 a = make_float2(0.f, 0.f);
  for (int z = 0; z < chunks; ++z) a = cadd(a, part[((long)z * (K - 1) + (k - 1)) * mm + q]);
  const float inv = (float)(1.0 / G);
  const float2 bp = make_float2((float)beta[0], (float)beta[1]);
  const long i = (long)k * mm + q;
  probe[i] = cadd(probe[i], cmul(cscale(a, inv), bp));
}

template <int M>
static hipError_t mode_updates_m(float2* probe, const float2* psi, const float2* op, const float* illum,
                                 const int2* roi, const int* gf, int K, int G, int cols, float fw,
                                 const double* beta, float2* part, int fpc, bool row_waves, const float2* tw,
                                 hipStream_t s) {
  using C = WaveCfg<M>;
  const int chunks = (G + fpc - 1) / fpc;
  bool launched = false;
  if constexpr (M <= 256) {
    if (row_waves) {
      hipLaunchKernelGGL((k_pum_waves<M, PUM_WAVES>), dim3(WaveTile<M>::T, K - 1, chunks), dim3(64 * PUM_WAVES), 0,
                         s, psi, op, illum, roi, gf, K, G, cols, fw, part, fpc, tw);
      launched = true;
    }
  }
  if (!launched)
    hipLaunchKernelGGL((k_pum_rows<M>), dim3(C::L / C::GPW, K - 1, chunks), dim3(PTYK_NT), 0, s, psi, op, illum, roi,
                       gf, K, G, cols, fw, part, fpc, tw);
  const long mm = (long)M * M;
  hipLaunchKernelGGL(k_pum_finish, dim3((unsigned)((mm + PTYK_NT - 1) / PTYK_NT), K - 1), dim3(PTYK_NT), 0, s, probe,
                     part, K, chunks, G, mm, beta);
  return hipGetLastError();
}

hipError_t mls_mode_updates(float2* probe, const float2* psi, const float2* op, const float* illum, const int2* roi,
                            const int* gf, int K, int G, int m, int cols, float fw, const double* beta, float2* part,
                            int fpc, bool row_waves, const float2* tw, hipStream_t s) {
  if (K < 2 || G < 1 || fpc < MLS_PUM_FPC_MIN) return hipErrorInvalidValue;
  switch (m) {
    case 128:
      return mode_updates_m<128>(probe, psi, op, illum, roi, gf, K, G, cols, fw, beta, part, fpc, row_waves, tw, s);
    case 256:
      return mode_updates_m<256>(probe, psi, op, illum, roi, gf, K, G, cols, fw, beta, part, fpc, row_waves, tw, s);
    case 512:
      return mode_updates_m<512>(probe, psi, op, illum, roi, gf, K, G, cols, fw, beta, part, fpc, row_waves, tw, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M>
static hipError_t mls_shift_m(float2* pu, const float2* chi, const float2* op, int G, const float2* sub,
                              const int* gf, float sign, const float2* tw, hipStream_t s) {
  constexpr int NB = plan_min_batch<M, PTYK_NT>(), NC = (NB + 15) / 16 * 16;
  hipLaunchKernelGGL((k_mls_pu_rows<M>), dim3((unsigned)((long)G * M / NB)), dim3(PTYK_NT), 0, s, pu, chi, op, sub, gf,
                     sign, tw);
  hipLaunchKernelGGL((k_mls_shift_cols<M>), dim3(M / NC, G), dim3(PTYK_NT), 0, s, pu, sub, gf, sign, tw);
  return hipGetLastError();
}

hipError_t mls_shift_cols(float2* pu, int G, int m, const float2* sub, const int* gf, float sign, const float2* tw,
                          hipStream_t s) {
  auto go = [&](auto kern, int nc) {
    hipLaunchKernelGGL(kern, dim3(m / nc, G), dim3(PTYK_NT), 0, s, pu, sub, gf, sign, tw);
    return hipGetLastError();
  };
  switch (m) {
    case 128: return go(k_mls_shift_cols<128>, (plan_min_batch<128, PTYK_NT>() + 15) / 16 * 16);
    case 256: return go(k_mls_shift_cols<256>, (plan_min_batch<256, PTYK_NT>() + 15) / 16 * 16);
    case 512: return go(k_mls_shift_cols<512>, (plan_min_batch<512, PTYK_NT>() + 15) / 16 * 16);
    default: return hipErrorInvalidValue;
  }
}

hipError_t mls_shift_passes(float2* pu, const float2* chi, const float2* op, int G, int m, const float2* sub,
                            const int* gf, float sign, const float2* tw, hipStream_t s) {
  switch (m) {
    case 128: return mls_shift_m<128>(pu, chi, op, G, sub, gf, sign, tw, s);
    case 256: return mls_shift_m<256>(pu, chi, op, G, sub, gf, sign, tw, s);
    case 512: return mls_shift_m<512>(pu, chi, op, G, sub, gf, sign, tw, s);
    default: return hipErrorInvalidValue;
  }
}

bool dm_wave_supported(int m) { return m == 128 || m == 256 || m == 512; }
int dm_wave_probe_blocks(int m) { return (m * m + PTYK_NT - 1) / PTYK_NT; }

hipError_t dm_wave_iteration(const DmWaveArgs& a, hipStream_t s, DmTimer* tm) {
  switch (a.m) {
    case 128: return wave_impl<128>(a, s, tm);
    case 256: return wave_impl<256>(a, s, tm);
    case 512: return wave_impl<512>(a, s, tm);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ptyk
