// Batched power-of-two 2D FFT over n frames of m x m complex64, staged in LDS.
//
// Replaces cufftPlanMany(rank 2, {H,W}) C2C plans (CudaFFTPlan.cpp:91-101) for the
// untangled call sites: probe initialisation (utilities.cpp:138-154), the simulator
// (Diffractions.cpp:105-107) and the FFT entry point used by the tests.
// Pass 1: each workgroup transforms a batch of whole rows (contiguous) in LDS.
// Pass 2: each workgroup transforms strips of 16 columns (128 B per row segment).
// Frames must start 16-byte aligned (every engine buffer does).
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

template <int M>
struct FftCfg {
  static constexpr int RNB = plan_min_batch<M, PTYK_NT>();
  static constexpr int CNB = (plan_min_batch<M, PTYK_NT>() + 15) / 16 * 16;
};

// Both passes move two adjacent complex values (16 B) per lane in fully unrolled loops, so every
// load of a thread is in flight before its first LDS store.  PAIR: two frames read from a and b
// (rows 0..M-1 and M..2M-1) and written to data.
template <int M, bool INV, bool PAIR = false>
__global__ __launch_bounds__(PTYK_NT) void k_fft_rows(float2* __restrict__ data, long total_rows,
                                                      const float2* __restrict__ tw,
                                                      const float2* __restrict__ a = nullptr,
                                                      const float2* __restrict__ b = nullptr) {
  constexpr int NB = FftCfg<M>::RNB;
  constexpr int PAIRS = NB * M / 2 / PTYK_NT;
  static_assert((NB * M / 2) % PTYK_NT == 0, "row tile must divide among the threads");
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * NB;
  {
    float4 v[PAIRS];
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      const long g = row0 + e / M;
      const float2* src = PAIR ? (g < M ? a + g * M : b + (g - M) * M) : data + g * M;
      v[i] = g < total_rows ? *reinterpret_cast<const float4*>(src + (e % M)) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
      const int e = 2 * (tid + i * PTYK_NT);
      *reinterpret_cast<float4*>(&lds[SwzRowLayout<M, M>::addr(e / M, e % M)]) = v[i];
    }
  }
  __syncthreads();
  fft_lds<M, INV, NB, PTYK_NT, SwzRowLayout<M, M>, false>(lds, tw, tid);
#pragma unroll
  for (int i = 0; i < PAIRS; ++i) {
    const int e = 2 * (tid + i * PTYK_NT);
    const long g = row0 + e / M;
    if (g < total_rows)
      *reinterpret_cast<float4*>(data + g * M + (e % M)) =
          *reinterpret_cast<const float4*>(&lds[SwzRowLayout<M, M>::addr(e / M, e % M)]);
  }
}

template <int M, bool INV>
__global__ __launch_bounds__(PTYK_NT) void k_fft_cols(float2* __restrict__ data, long total_strips,
                                                      const float2* __restrict__ tw) {
  constexpr int NB = FftCfg<M>::CNB;
  constexpr int G = NB / 16;
  constexpr int SPF = M / 16;
  constexpr int PAIRS = NB * M / 2 / PTYK_NT;
  static_assert((NB * M / 2) % PTYK_NT == 0, "column tile must divide among the threads");
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  {
    float4 v[PAIRS];
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) {
      const int e = 2 * (tid + k * PTYK_NT);
      const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
      const long s = (long)blockIdx.x * G + g;
      v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (s < total_strips) {
        const long f = s / SPF;
        const int col = (int)(s % SPF) * 16 + c;
        v[k] = *reinterpret_cast<const float4*>(data + (f * M + i) * M + col);
      }
    }
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) *reinterpret_cast<float4*>(&lds[2 * (tid + k * PTYK_NT)]) = v[k];
  }
  __syncthreads();
  fft_lds<M, INV, NB, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
#pragma unroll
  for (int k = 0; k < PAIRS; ++k) {
    const int e = 2 * (tid + k * PTYK_NT);
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    const long s = (long)blockIdx.x * G + g;
    if (s < total_strips) {
      const long f = s / SPF;
      const int col = (int)(s % SPF) * 16 + c;
      *reinterpret_cast<float4*>(data + (f * M + i) * M + col) = *reinterpret_cast<const float4*>(&lds[e]);
    }
  }
}

template <int M, bool INV>
static hipError_t fft2_impl(float2* data, long n, const float2* tw, hipStream_t s) {
  const long rows = n * M;
  const long strips = n * (M / 16);
  const long gr = (rows + FftCfg<M>::RNB - 1) / FftCfg<M>::RNB;
  const long gc = (strips + FftCfg<M>::CNB / 16 - 1) / (FftCfg<M>::CNB / 16);
  if (gr > 0x7fffffffL || gc > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_fft_rows<M, INV>), dim3((unsigned)gr), dim3(PTYK_NT), 0, s, data, rows, tw);
  hipLaunchKernelGGL((k_fft_cols<M, INV>), dim3((unsigned)gc), dim3(PTYK_NT), 0, s, data, strips, tw);
  return hipGetLastError();
}

template <int M, bool INV>
static hipError_t fft_cols_impl(float2* data, long n, const float2* tw, hipStream_t s) {
  const long strips = n * (M / 16);
  const long gc = (strips + FftCfg<M>::CNB / 16 - 1) / (FftCfg<M>::CNB / 16);
  if (gc > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_fft_cols<M, INV>), dim3((unsigned)gc), dim3(PTYK_NT), 0, s, data, strips, tw);
  return hipGetLastError();
}

template <int M>
static hipError_t fft2_pair_impl(float2* dst, const float2* a, const float2* b, const float2* tw, hipStream_t s) {
  const long rows = 2L * M, strips = 2L * (M / 16);
  const long gr = (rows + FftCfg<M>::RNB - 1) / FftCfg<M>::RNB;
  const long gc = (strips + FftCfg<M>::CNB / 16 - 1) / (FftCfg<M>::CNB / 16);
  hipLaunchKernelGGL((k_fft_rows<M, false, true>), dim3((unsigned)gr), dim3(PTYK_NT), 0, s, dst, rows, tw, a, b);
  hipLaunchKernelGGL((k_fft_cols<M, false>), dim3((unsigned)gc), dim3(PTYK_NT), 0, s, dst, strips, tw);
  return hipGetLastError();
}

template <int M>
static hipError_t fft_rows_pair_impl(float2* dst, const float2* a, const float2* b, const float2* tw, hipStream_t s) {
  const long rows = 2L * M;
  const long gr = (rows + FftCfg<M>::RNB - 1) / FftCfg<M>::RNB;
  hipLaunchKernelGGL((k_fft_rows<M, false, true>), dim3((unsigned)gr), dim3(PTYK_NT), 0, s, dst, rows, tw, a, b);
  return hipGetLastError();
}

hipError_t fft_rows_pair(float2* dst, const float2* a, const float2* b, int m, const float2* tw, hipStream_t s) {
  switch (m) {
    case 128: return fft_rows_pair_impl<128>(dst, a, b, tw, s);
    case 256: return fft_rows_pair_impl<256>(dst, a, b, tw, s);
    case 512: return fft_rows_pair_impl<512>(dst, a, b, tw, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft2_pair(float2* dst, const float2* a, const float2* b, int m, const float2* tw, hipStream_t s) {
  switch (m) {
    case 16: return fft2_pair_impl<16>(dst, a, b, tw, s);
    case 32: return fft2_pair_impl<32>(dst, a, b, tw, s);
    case 64: return fft2_pair_impl<64>(dst, a, b, tw, s);
    case 128: return fft2_pair_impl<128>(dst, a, b, tw, s);
    case 256: return fft2_pair_impl<256>(dst, a, b, tw, s);
    case 512: return fft2_pair_impl<512>(dst, a, b, tw, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft_cols_batch(float2* data, long n, int m, bool inv, const float2* tw, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  switch (m) {
    case 128: return inv ? fft_cols_impl<128, true>(data, n, tw, s) : fft_cols_impl<128, false>(data, n, tw, s);
    case 256: return inv ? fft_cols_impl<256, true>(data, n, tw, s) : fft_cols_impl<256, false>(data, n, tw, s);
    case 512: return inv ? fft_cols_impl<512, true>(data, n, tw, s) : fft_cols_impl<512, false>(data, n, tw, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft2_batch(float2* data, long n, int m, bool inv, const float2* tw, hipStream_t s) {
  if (n <= 0) return hipSuccess;
#define PTYK_FFT_CASE(MM) \
  case MM: return inv ? fft2_impl<MM, true>(data, n, tw, s) : fft2_impl<MM, false>(data, n, tw, s);
  switch (m) {
    PTYK_FFT_CASE(16)
    PTYK_FFT_CASE(32)
    PTYK_FFT_CASE(64)
    PTYK_FFT_CASE(128)
    PTYK_FFT_CASE(256)
    PTYK_FFT_CASE(512)
    default: return hipErrorInvalidValue;
  }
#undef PTYK_FFT_CASE
}

}  // namespace ptyk
