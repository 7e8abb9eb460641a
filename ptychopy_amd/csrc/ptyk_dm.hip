// Difference-map iteration on gfx950 (replaces DM.cu:80-128 and the cuFFT load/store
// callbacks d_calculatePsi / d_modulusConstraint / d_updateObject, SampleKernels.cu:418-488).
//
// Reference equations per frame j, ROI origin r_j = roundf(pos_j) (SampleKernels.cu:426,460):
//   x_j   = 2 P O_r - psi_j
//   Phi   = FFT(x_j);  Phi' = A_j exp(i arg Phi) / M^2           (d_modulusConstraint)
//   Pf    = IFFT_u(Phi')
//   psi_j <- psi_j + Pf - 2 P O_r                                   (d_updateObject)
//   O_r  += s_O conj(P) psi_j   (phase-clamped increment when the flag is set)
//   P    += s_P conj(O_r) psi_j (when updateProbe)
// with s_O = 1/probeMax and s_P = 1/objMax taken from the cached factors (DM.cu:46-73).
// Object and probe reads use the values at the start of the iteration and all frame
// increments are summed in a fixed order (no float atomics).
//
// Data flow (ψ state stays in one N x M x M buffer, rewritten in place):
//   k_dm_rows_fwd : X = RowFFT(2 P O_r - psi)                          (rows)
//   k_dm_cols     : Z = ColIFFT_u((A/|Phi| - 1) Phi / M^2),  Phi = ColFFT(X)
//                   -- the linear identity psi_new = Pf - x = IFFT2_u(Phi' - Phi/M^2)
//                      removes the need to keep psi_old or X.
//   k_dm_rows_inv : psi_new = RowIFFT_u(Z)
//   k_dm_probe_partial : probe increment partial sums over frame chunks (reads O at start)
//   k_dm_object   : object increment gathered per object row over the frames that cover
//                   it (CSR list, ascending frame id; reads P at start, rewrites O)
//   k_dm_probe_apply : P += s_P * sum of partials (fixed chunk order)
//   k_dm_finish   : FP64 sums -> objMax, probeMax; cached factor update (|delta| > 1e-3)
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

template <int M>
struct DmCfg {
  static constexpr int RNB = plan_min_batch<M, PTYK_NT>();
  static constexpr int CNB = (plan_min_batch<M, PTYK_NT>() + 15) / 16 * 16;
  static constexpr int PCPT = M >= PTYK_NT ? M / PTYK_NT : 1;   // probe columns per thread
  static constexpr int PRPW = M >= PTYK_NT ? 1 : PTYK_NT / M;   // probe rows per workgroup
};

#define PTYK_OBJ_CHUNK 256

template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dm_rows_fwd(float2* __restrict__ psi, const float2* __restrict__ obj,
                                                         int obj_cols, const float2* __restrict__ probe,
                                                         const int2* __restrict__ roi, long total_rows,
                                                         const float2* __restrict__ tw) {
  constexpr int NB = DmCfg<M>::RNB;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * NB;
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const int c = e % M;
    const long g = row0 + e / M;
    float2 v = make_float2(0.f, 0.f);
    if (g < total_rows) {
      const long f = g / M;
      const int i = (int)(g - f * M);
      const int2 r = roi[f];
      const float2 po = cmul(probe[i * M + c], obj[(long)(r.x + i) * obj_cols + (r.y + c)]);
      const float2 s = psi[g * M + c];
      v = make_float2(2.f * po.x - s.x, 2.f * po.y - s.y);
    }
    lds[e] = v;
  }
  __syncthreads();
  fft_lds<M, false, NB, PTYK_NT, RowLayout<M, M>, false>(lds, tw, tid);
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const long g = row0 + e / M;
    if (g < total_rows) psi[g * M + (e % M)] = lds[e];
  }
}

template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dm_cols(float2* __restrict__ psi, const __half* __restrict__ amp,
                                                     const float* __restrict__ amp_scale, long total_strips,
                                                     const float2* __restrict__ tw) {
  constexpr int NB = DmCfg<M>::CNB;
  constexpr int G = NB / 16;
  constexpr int SPF = M / 16;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    const long s = (long)blockIdx.x * G + g;
    float2 v = make_float2(0.f, 0.f);
    if (s < total_strips) {
      const long f = s / SPF;
      v = psi[(f * M + i) * M + (int)(s % SPF) * 16 + c];
    }
    lds[e] = v;
  }
  __syncthreads();
  fft_lds<M, false, NB, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
  // Fourier modulus constraint, written as the difference Phi' - Phi/M^2.
  const float inv_m2 = 1.f / (float)(M * M);
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    const long s = (long)blockIdx.x * G + g;
    float2 d = make_float2(0.f, 0.f);
    if (s < total_strips) {
      const long f = s / SPF;
      const float a = load_amp(amp, (f * M + i) * M + (int)(s % SPF) * 16 + c, amp_scale[f]);
      const float2 ph = lds[e];
      const float m2 = cabs2(ph);
      if (m2 > 0.f) {
        const float k = (a * rsqrtf(m2) - 1.f) * inv_m2;
        d = make_float2(ph.x * k, ph.y * k);
      } else {
        d = make_float2(a * inv_m2, 0.f);  // atan2(0,0) = 0 in the reference
      }
    }
    lds[e] = d;
  }
  __syncthreads();
  fft_lds<M, true, NB, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    const long s = (long)blockIdx.x * G + g;
    if (s < total_strips) {
      const long f = s / SPF;
      psi[(f * M + i) * M + (int)(s % SPF) * 16 + c] = lds[e];
    }
  }
}

template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dm_rows_inv(float2* __restrict__ psi, long total_rows,
                                                         const float2* __restrict__ tw) {
  constexpr int NB = DmCfg<M>::RNB;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * NB;
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const long g = row0 + e / M;
    lds[e] = g < total_rows ? psi[g * M + (e % M)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<M, true, NB, PTYK_NT, RowLayout<M, M>, false>(lds, tw, tid);
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const long g = row0 + e / M;
    if (g < total_rows) psi[g * M + (e % M)] = lds[e];
  }
}

// partial[c][q][col] = sum over frames j in chunk c of conj(O[r_j + (q,col)]) psi_j[q][col]
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dm_probe_partial(const float2* __restrict__ psi,
                                                              const float2* __restrict__ obj, int obj_cols,
                                                              const int2* __restrict__ roi, long n_frames,
                                                              int frames_per_chunk, double2* __restrict__ partial) {
  constexpr int CPT = DmCfg<M>::PCPT;
  constexpr int RPW = DmCfg<M>::PRPW;
  const int q = blockIdx.x * RPW + (RPW > 1 ? threadIdx.x / M : 0);
  const int c0 = RPW > 1 ? threadIdx.x % M : threadIdx.x;
  const long j0 = (long)blockIdx.y * frames_per_chunk;
  const long j1 = j0 + frames_per_chunk < n_frames ? j0 + frames_per_chunk : n_frames;
  double ax[CPT], ay[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) ax[k] = ay[k] = 0.0;
  for (long j = j0; j < j1; ++j) {
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
    partial[((long)blockIdx.y * M + q) * M + c] = make_double2(ax[k], ay[k]);
  }
}

// P += s_P * sum_c partial[c] (fixed chunk order) when update; always returns sum |P|^2 per block.
template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dm_probe_apply(float2* __restrict__ probe,
                                                            const double2* __restrict__ partial, int chunks,
                                                            const float* __restrict__ factors, int update,
                                                            double* __restrict__ block_sums) {
  constexpr int CPT = DmCfg<M>::PCPT;
  constexpr int RPW = DmCfg<M>::PRPW;
  __shared__ double red[PTYK_NT / 64];
  const int q = blockIdx.x * RPW + (RPW > 1 ? threadIdx.x / M : 0);
  const int c0 = RPW > 1 ? threadIdx.x % M : threadIdx.x;
  const float sp = factors[0];
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int c = c0 + k * PTYK_NT;
    float2 p = probe[q * M + c];
    if (update) {
      double sx = 0.0, sy = 0.0;
      for (int ch = 0; ch < chunks; ++ch) {
        const double2 t = partial[((long)ch * M + q) * M + c];
        sx += t.x;
        sy += t.y;
      }
      p.x = (float)((double)p.x + (double)sp * sx);
      p.y = (float)((double)p.y + (double)sp * sy);
      probe[q * M + c] = p;
    }
    acc += (double)p.x * p.x + (double)p.y * p.y;
  }
  acc = block_sum<PTYK_NT>(acc, red);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = acc;
}

template <int M>
__global__ __launch_bounds__(PTYK_NT) void k_dm_object(float2* __restrict__ obj, int obj_cols,
                                                       const float2* __restrict__ psi,
                                                       const float2* __restrict__ probe,
                                                       const int2* __restrict__ roi,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ row_frames,
                                                       const float* __restrict__ factors, int phase_constraint,
                                                       double* __restrict__ block_sums) {
  __shared__ double red[PTYK_NT / 64];
  const int p = blockIdx.x;
  const int c = blockIdx.y * PTYK_OBJ_CHUNK + threadIdx.x;
  const int cbeg = blockIdx.y * PTYK_OBJ_CHUNK;
  const int cend = cbeg + PTYK_OBJ_CHUNK;
  const float so = factors[1];
  double ax = 0.0, ay = 0.0;
  const int k0 = row_ptr[p], k1 = row_ptr[p + 1];
  for (int k = k0; k < k1; ++k) {
    const int j = row_frames[k];
    const int2 r = roi[j];
    if (r.y >= cend || r.y + M <= cbeg) continue;  // frame does not touch this column chunk
    const int q = p - r.x;
    const int cc = c - r.y;
    if (c < obj_cols && cc >= 0 && cc < M) {
      float2 u = cmulcj(probe[q * M + cc], psi[((long)j * M + q) * M + cc]);
      u = cscale(u, so);
      if (phase_constraint) u = clamp_phase(u);
      ax += u.x;
      ay += u.y;
    }
  }
  double s2 = 0.0;
  if (c < obj_cols) {
    float2 o = obj[(long)p * obj_cols + c];
    o.x = (float)((double)o.x + ax);
    o.y = (float)((double)o.y + ay);
    obj[(long)p * obj_cols + c] = o;
    s2 = (double)o.x * o.x + (double)o.y * o.y;
  }
  s2 = block_sum<PTYK_NT>(s2, red);
  if (threadIdx.x == 0) block_sums[(long)blockIdx.x * gridDim.y + blockIdx.y] = s2;
}

int dm_obj_col_chunks(int obj_cols) { return (obj_cols + PTYK_OBJ_CHUNK - 1) / PTYK_OBJ_CHUNK; }

template <int M>
static hipError_t dm_impl(const DmArgs& a, hipStream_t s, DmTimer* tm) {
  const long rows = a.n_frames * M;
  const long strips = a.n_frames * (M / 16);
  const unsigned gr = (unsigned)((rows + DmCfg<M>::RNB - 1) / DmCfg<M>::RNB);
  const unsigned gc = (unsigned)((strips + DmCfg<M>::CNB / 16 - 1) / (DmCfg<M>::CNB / 16));
  const unsigned gp = (unsigned)(M / DmCfg<M>::PRPW);
  auto rec = [&](int cat) {
    if (tm) tm->mark(s, cat);
  };
  rec(DM_T_ROWS_FWD);
  hipLaunchKernelGGL((k_dm_rows_fwd<M>), dim3(gr), dim3(PTYK_NT), 0, s, a.psi, a.obj, a.obj_cols, a.probe, a.roi,
                     rows, a.tw);
  rec(DM_T_COLS);
  hipLaunchKernelGGL((k_dm_cols<M>), dim3(gc), dim3(PTYK_NT), 0, s, a.psi, a.amp, a.amp_scale, strips, a.tw);
  rec(DM_T_ROWS_INV);
  hipLaunchKernelGGL((k_dm_rows_inv<M>), dim3(gr), dim3(PTYK_NT), 0, s, a.psi, rows, a.tw);
  if (a.update_probe) {
    rec(DM_T_PROBE_GATHER);
    hipLaunchKernelGGL((k_dm_probe_partial<M>), dim3(gp, a.probe_chunks), dim3(PTYK_NT), 0, s, a.psi, a.obj,
                       a.obj_cols, a.roi, a.n_frames, a.frames_per_chunk, a.probe_partial);
  }
  const int cch = dm_obj_col_chunks(a.obj_cols);
  rec(DM_T_OBJ_GATHER);
  hipLaunchKernelGGL((k_dm_object<M>), dim3(a.obj_rows, cch), dim3(PTYK_NT), 0, s, a.obj, a.obj_cols, a.psi,
                     a.probe, a.roi, a.row_ptr, a.row_frames, a.factors, a.phase_constraint, a.obj_partial);
  rec(DM_T_PROBE_APPLY);
  hipLaunchKernelGGL((k_dm_probe_apply<M>), dim3(gp), dim3(PTYK_NT), 0, s, a.probe, a.probe_partial,
                     a.probe_chunks, a.factors, a.update_probe, a.probe_partial_sum);
  rec(DM_T_FINISH);
  hipLaunchKernelGGL((k_dm_finish<PTYK_NT>), dim3(1), dim3(PTYK_NT), 0, s, a.obj_partial, a.obj_rows * cch,
                     a.probe_partial_sum, (int)gp, a.scal, a.factors);
  rec(-1);
  return hipGetLastError();
}

hipError_t dm_iteration(const DmArgs& a, hipStream_t s, DmTimer* ev) {
  switch (a.m) {
    case 16: return dm_impl<16>(a, s, ev);
    case 32: return dm_impl<32>(a, s, ev);
    case 64: return dm_impl<64>(a, s, ev);
    case 128: return dm_impl<128>(a, s, ev);
    case 256: return dm_impl<256>(a, s, ev);
    case 512: return dm_impl<512>(a, s, ev);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ptyk
