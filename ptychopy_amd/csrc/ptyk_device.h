// Device building blocks for the gfx950 ptychography engine.
//
//  * complex float2 arithmetic,
//  * in-register radix-2/4/8/16 DFTs,
//  * a compile-time Stockham autosort plan that runs a batch of power-of-two
//    transforms through LDS (replaces the cuFFT plans of CudaFFTPlan.cpp:63-101),
//  * wave-shuffle reductions with FP64 accumulation (replaces reductions.cu).
//
// Conventions: forward transform uses exp(-2*pi*i*k*n/N); the inverse is
// unnormalised, as cuFFT's CUFFT_INVERSE is (CudaFFTPlan.h:89-125).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define PTYK_NT 256  // threads per workgroup of the batched frame kernels

namespace ptyk {

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ float2 cmulcj(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 cmul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
__device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }

// A exp(i arg f) * nrm with arg computed as atan2f (d_adjustFFT's modulusConstraint,
// utilitiesKernels.cu:626-631); a zero f takes phase 0 whatever the signs of its zeros.
__device__ __forceinline__ float2 modulus_phase(float2 f, float a, float nrm) {
  if (f.x == 0.f && f.y == 0.f) return make_float2(a * nrm, 0.f);
  float sn, cs;
  sincosf(atan2f(f.y, f.x), &sn, &cs);
  return make_float2(a * cs * nrm, a * sn * nrm);
}

// Reference phase clamp: |u| * exp(i * min(arg u, 0))  (SampleKernels.cu:209-215)
__device__ __forceinline__ float2 clamp_phase(float2 u) {
  float mag = sqrtf(u.x * u.x + u.y * u.y);
  float arg = atan2f(u.y, u.x);
  arg = arg <= 0.f ? arg : 0.f;
  float s, c;
  sincosf(arg, &s, &c);
  return make_float2(mag * c, mag * s);
}

// ---------------------------------------------------------------- small DFTs
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct Dft<4, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <bool INV>
struct Dft<8, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    const float h = 0.70710678118654752f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    // twiddles W8^k, k = 0..3 (forward: exp(-i*pi*k/4))
    float2 o1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                    : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    float2 o2 = cmul_mi<INV>(o[2]);
    float2 o3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                    : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <bool INV>
struct Dft<16, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    // 16 = 4 x 4: inner DFT4 over n1 (stride 4), twiddle W16^(n2*k1), outer DFT4 over n2.
    const float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, h = 0.70710678118654752f;
    const float sg = INV ? 1.f : -1.f;
    float2 b[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      float2 t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<4, INV>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) b[n2][k1] = t[k1];
    }
    // twiddles W16^(n2*k1) for n2,k1 in 1..3: exponents 1,2,3,2,4,6,3,6,9
    const float cw[10] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1};
    const float sw[10] = {0.f, s1, h, c1, 1.f, c1, h, s1, 0.f, -s1};
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) {
        int e = n2 * k1;
        float2 w = make_float2(cw[e], sg * sw[e]);
        b[n2][k1] = cmul(b[n2][k1], w);
      }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 t[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
      Dft<4, INV>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// ---------------------------------------------------------------- FFT plans
// Radix sequence per transform length; Ns (the sub-transform length before a
// stage) is the product of the previous radices.
template <int N>
struct Plan;
template <> struct Plan<16>  { static constexpr int S = 2; static constexpr int R0 = 4, R1 = 4, R2 = 1; };
template <> struct Plan<32>  { static constexpr int S = 2; static constexpr int R0 = 4, R1 = 8, R2 = 1; };
template <> struct Plan<64>  { static constexpr int S = 2; static constexpr int R0 = 8, R1 = 8, R2 = 1; };
template <> struct Plan<128> { static constexpr int S = 2; static constexpr int R0 = 8, R1 = 16, R2 = 1; };
template <> struct Plan<256> { static constexpr int S = 2; static constexpr int R0 = 16, R1 = 16, R2 = 1; };
template <> struct Plan<512> { static constexpr int S = 3; static constexpr int R0 = 8, R1 = 8, R2 = 8; };

template <int N>
__host__ __device__ constexpr int plan_radix(int s) {
  return s == 0 ? Plan<N>::R0 : (s == 1 ? Plan<N>::R1 : Plan<N>::R2);
}
template <int N>
__host__ __device__ constexpr int plan_ns(int s) {
  return s == 0 ? 1 : plan_ns<N>(s - 1) * plan_radix<N>(s - 1);
}
template <int N>
__host__ __device__ constexpr int plan_rmax() {
  return Plan<N>::R0 > Plan<N>::R1 ? (Plan<N>::R0 > Plan<N>::R2 ? Plan<N>::R0 : Plan<N>::R2)
                                   : (Plan<N>::R1 > Plan<N>::R2 ? Plan<N>::R1 : Plan<N>::R2);
}
// Smallest batch count so that every stage has a whole number of butterflies per thread.
template <int N, int NT>
__host__ __device__ constexpr int plan_min_batch() {
  return (NT * plan_rmax<N>()) / N > 0 ? (NT * plan_rmax<N>()) / N : 1;
}

// Twiddle table: tw[k] = exp(-2*pi*i*k/TWN), k in [0, TWN), computed on the host in FP64.
#define PTYK_TWN 1024

// Layouts of a batch of NB sequences of length N inside LDS.
// Rows: sequence b is contiguous, stride BS between sequences.
template <int N, int BS>
struct RowLayout {
  __device__ __forceinline__ static int addr(int b, int e) { return b * BS + e; }
};
// Rows whose elements are permuted to spread LDS banks: element e of a row of N >= 256 sits at
// row_swz<N>(e) = e ^ (((e >> 1) ^ (e >> 3)) & 30), a bijection within each aligned block of 32
// that leaves bit 0 alone (so the pairs (2i, 2i + 1) stay adjacent, in order and 16-byte
// aligned). Unpermuted, the first Stockham stage's writes (stride R between lanes) meet 8-way
// (N = 512) and 16-way (N = 256) bank conflicts; permuted, no access of the row transforms
// exceeds 2-way, for 32 or 64 banks of 4 bytes.
template <int N>
__device__ __forceinline__ int row_swz(int e) {
  if constexpr (N >= 256)
    return e ^ (((e >> 1) ^ (e >> 3)) & 30);
  else
    return e;
}
template <int N, int BS>
struct SwzRowLayout {
  __device__ __forceinline__ static int addr(int b, int e) { return b * BS + row_swz<N>(e); }
};
// Column strips: 16 columns per strip, strip-major, column index fastest.
template <int N>
struct StripLayout {
  __device__ __forceinline__ static int addr(int b, int e) { return (b & 15) + (b >> 4) * (N * 16) + e * 16; }
};

// Column block of one frame: LDS holds rows e = 0..N-1 of PITCH adjacent columns b.
template <int PITCH>
struct ColLayout {
  __device__ __forceinline__ static int addr(int b, int e) { return e * PITCH + b; }
};

// LDS exchange point.  Workgroup form: a barrier.  Wave form: the wave's LDS operations issued
// so far complete (lgkmcnt(0); the LDS serves one wave's operations in issue order) and the
// compiler may move no LDS access across the point; no other wave is waited for.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // gfx9 encoding: lgkmcnt(0), vmcnt and expcnt not waited on
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One radix-R Stockham stage over NB sequences, NT threads, in-place through registers.
// BATCH_FAST: consecutive threads take consecutive sequences (strip layout) rather than
// consecutive butterflies of one sequence (row layout).
// PREFETCH: issue every twiddle load of the thread before the first use (interleaved with the
// LDS reads and the products, the scheduler waited on each load in turn); off where the extra
// registers would spill (1024-thread workgroups).
// WAVE: the NT = 64 threads of one wave own the sequences, so the exchanges between the stages
// need only the wave's own LDS operations to complete (lds_sync<true>), not a workgroup barrier:
// the other waves of the workgroup run their own transforms independently.
template <int N, int R, int NS, bool INV, int NB, int NT, class L, bool BATCH_FAST, bool PREFETCH = true,
          bool WAVE = false>
__device__ __forceinline__ void stockham_stage(float2* lds, const float2* __restrict__ tw, int tid) {
  static_assert(!WAVE || NT == 64, "a wave-local transform runs on one wave");
  constexpr int NBF = NB * (N / R);
  static_assert(NBF % NT == 0, "butterflies must divide evenly among threads");
  constexpr int BPT = NBF / NT;
  constexpr int TWS = PTYK_TWN / (NS * R);  // twiddle table stride for this stage
  float2 v[BPT][R];
  int bb[BPT], jj[BPT];
  float2 w[BPT][R > 1 ? R : 2];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int t = tid + q * NT;
    bb[q] = BATCH_FAST ? (t % NB) : (t / (N / R));
    jj[q] = BATCH_FAST ? (t / NB) : (t % (N / R));
    if (PREFETCH && NS > 1) {
      const int k = jj[q] & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) w[q][r] = tw[(k * r * TWS) & (PTYK_TWN - 1)];
    }
  }
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = lds[L::addr(bb[q], jj[q] + r * (N / R))];
  }
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    if (NS > 1) {
      const int k = jj[q] & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 wr = PREFETCH ? w[q][r] : tw[(k * r * TWS) & (PTYK_TWN - 1)];
        if (INV) wr.y = -wr.y;
        v[q][r] = cmul(v[q][r], wr);
      }
    }
    Dft<R, INV>::run(v[q]);
  }
  lds_sync<WAVE>();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int j = jj[q];
    const int idx = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
    for (int r = 0; r < R; ++r) lds[L::addr(bb[q], idx + r * NS)] = v[q][r];
  }
  lds_sync<WAVE>();
}

// Full transform of NB sequences already in LDS (caller syncs before the call).
template <int N, bool INV, int NB, int NT, class L, bool BATCH_FAST, bool PREFETCH = true, bool WAVE = false>
__device__ __forceinline__ void fft_lds(float2* lds, const float2* __restrict__ tw, int tid) {
  stockham_stage<N, plan_radix<N>(0), plan_ns<N>(0), INV, NB, NT, L, BATCH_FAST, PREFETCH, WAVE>(lds, tw, tid);
  if constexpr (Plan<N>::S > 1)
    stockham_stage<N, plan_radix<N>(1), plan_ns<N>(1), INV, NB, NT, L, BATCH_FAST, PREFETCH, WAVE>(lds, tw, tid);
  if constexpr (Plan<N>::S > 2)
    stockham_stage<N, plan_radix<N>(2), plan_ns<N>(2), INV, NB, NT, L, BATCH_FAST, PREFETCH, WAVE>(lds, tw, tid);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
// Block sum with a fixed combination order (deterministic); every thread gets the result.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* scratch /* NT/64 doubles */) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) s += scratch[i];
  __syncthreads();
  return s;
}
template <int NT>
__device__ __forceinline__ float block_max(float v, float* scratch /* NT/64 floats */) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[w] = v;
  __syncthreads();
  float s = scratch[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) s = fmaxf(s, scratch[i]);
  __syncthreads();
  return s;
}

// DM iteration epilogue: FP64 totals of the per-block partial sums in a fixed order, then the
// reference's cached-factor rule (DM::checkCache, DM.cu:54-73): a cached factor is replaced
// only when it moved by more than 1e-3.  factors = {1/objMax, 1/probeMax}.
template <int NT>
__global__ __launch_bounds__(NT) void k_dm_finish(const double* __restrict__ obj_partial, int n_obj,
                                                  const double* __restrict__ probe_partial, int n_probe,
                                                  double* __restrict__ scal, float* __restrict__ factors) {
  __shared__ double red[NT / 64];
  double so = 0.0, sp = 0.0;
  for (int i = threadIdx.x; i < n_obj; i += NT) so += obj_partial[i];
  for (int i = threadIdx.x; i < n_probe; i += NT) sp += probe_partial[i];
  so = block_sum<NT>(so, red);
  sp = block_sum<NT>(sp, red);
  if (threadIdx.x == 0) {
    scal[0] = so;
    scal[1] = sp;
    const float passed0 = (float)(1.0 / so), passed1 = (float)(1.0 / sp);
    if (fabsf(factors[0] - passed0) > 1e-3f) factors[0] = passed0;
    if (fabsf(factors[1] - passed1) > 1e-3f) factors[1] = passed1;
  }
}

// FP16 amplitude with a per-frame FP32 scale.
__device__ __forceinline__ float load_amp(const __half* a, long idx, float scale) {
  return __half2float(a[idx]) * scale;
}

}  // namespace ptyk
