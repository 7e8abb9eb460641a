// ePIE iteration on gfx950 (replaces ePIE.cpp:65-129 and the kernels it drives:
// d_extractObjectArray (SampleKernels.cu:167-178), d_complexMultiply, the cuFFT pair of
// applyModulusConstraint (utilities.cpp:138-154) with d_adjustFFT / d_adjustModalFFT
// (utilitiesKernels.cu:633-682), d_updateObjectArray(M) (SampleKernels.cu:244-294),
// d_updateProbe(M) (ProbeKernels.cu:425-458), the thrust maxima behind
// Sample/Probe::updateMaxIntensity and d_calculateER (utilitiesKernels.cu:684-712)).
//
// ePIE is a chain of dependent position updates, so one persistent workgroup walks every
// position of an iteration in the given order: one launch per iteration and no host round
// trip per position. Single-mode frames of 64^2 and 128^2 keep the whole exit wave in LDS
// (k_epie_lds); other sizes and the modal form keep the exit waves in an L2-resident scratch
// and stage each 2D transform through LDS in row and column chunks (k_epie).
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

#define PTYK_EPIE_NT 1024

template <int M, int NT>
struct WgFft {
  static constexpr int RNB = plan_min_batch<M, NT>();
  static constexpr int CNB = (plan_min_batch<M, NT>() + 15) / 16 * 16;
  static constexpr int LDS = (RNB > CNB ? RNB : CNB) * M;  // float2 elements
};

// In-place 2D transform of one M x M frame held in global memory by one workgroup.
template <int M, bool INV, int NT>
__device__ void wg_fft2(float2* __restrict__ g, float2* lds, const float2* __restrict__ tw) {
  constexpr int RNB = WgFft<M, NT>::RNB;
  constexpr int CNB = WgFft<M, NT>::CNB;
  const int tid = threadIdx.x;
  for (int r0 = 0; r0 < M; r0 += RNB) {
    for (int e = tid; e < RNB * M; e += NT) {
      const int row = r0 + e / M;
      lds[e] = row < M ? g[row * M + (e % M)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<M, INV, RNB, NT, RowLayout<M, M>, false, false>(lds, tw, tid);
    for (int e = tid; e < RNB * M; e += NT) {
      const int row = r0 + e / M;
      if (row < M) g[row * M + (e % M)] = lds[e];
    }
    __syncthreads();
  }
  for (int c0 = 0; c0 < M; c0 += CNB) {
    for (int e = tid; e < CNB * M; e += NT) {
      const int col = c0 + e % CNB;
      lds[e] = col < M ? g[(e / CNB) * M + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<M, INV, CNB, NT, ColLayout<CNB>, true, false>(lds, tw, tid);
    for (int e = tid; e < CNB * M; e += NT) {
      const int col = c0 + e % CNB;
      if (col < M) g[(e / CNB) * M + col] = lds[e];
    }
    __syncthreads();
  }
}

// Bilinear sample with clamp addressing, matching a linear-filtered, unnormalised texture
// fetch at (u, v) = (col + offY + 0.5, row + offX + 0.5) (SampleKernels.cu:167-178, 493-508).
// quantize: round the weights to 1/256 (8 fractional bits, as texture units store them).
// The sample in steps (tap position and weights; clamped tap indices; the four gathers; the
// weighted sum), so that callers can issue the gathers of several samples before combining any,
// or take the taps from a staged copy of the object.
struct TapPos {
  int i0, j0;  // column and row of the first tap, before clamping
  float al, be;
};
template <bool QUANT>
__device__ __forceinline__ TapPos tap_pos(float u, float v) {
  const float xb = u - 0.5f, yb = v - 0.5f;
  const float fx = floorf(xb), fy = floorf(yb);
  TapPos p;
  p.al = xb - fx;
  p.be = yb - fy;
  if (QUANT) {
    p.al = rintf(p.al * 256.f) * (1.f / 256.f);
    p.be = rintf(p.be * 256.f) * (1.f / 256.f);
  }
  p.i0 = (int)fx;
  p.j0 = (int)fy;
  return p;
}
__device__ __forceinline__ float2 bilinear_sum(float al, float be, float2 t00, float2 t01, float2 t10, float2 t11) {
  const float w00 = (1.f - al) * (1.f - be), w01 = al * (1.f - be), w10 = (1.f - al) * be, w11 = al * be;
  return make_float2(w00 * t00.x + w01 * t01.x + w10 * t10.x + w11 * t11.x,
                     w00 * t00.y + w01 * t01.y + w10 * t10.y + w11 * t11.y);
}
template <bool QUANT>
__device__ __forceinline__ float2 bilinear_t(const float2* __restrict__ o, int rows, int cols, float u, float v) {
  const TapPos p = tap_pos<QUANT>(u, v);
  const int i0 = min(max(p.i0, 0), cols - 1), i1 = min(max(p.i0 + 1, 0), cols - 1);
  const int j0 = min(max(p.j0, 0), rows - 1), j1 = min(max(p.j0 + 1, 0), rows - 1);
  return bilinear_sum(p.al, p.be, o[(long)j0 * cols + i0], o[(long)j0 * cols + i1], o[(long)j1 * cols + i0],
                      o[(long)j1 * cols + i1]);
}
__device__ __forceinline__ float2 bilinear(const float2* __restrict__ o, int rows, int cols, float u, float v,
                                           int quantize) {
  return quantize ? bilinear_t<true>(o, rows, cols, u, v) : bilinear_t<false>(o, rows, cols, u, v);
}

template <int M>
__global__ __launch_bounds__(PTYK_EPIE_NT) void k_epie(EpieArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float redf[PTYK_EPIE_NT / 64];
  __shared__ double redd[PTYK_EPIE_NT / 64];
  constexpr int NT = PTYK_EPIE_NT;
  constexpr int MM = M * M;
  const int tid = threadIdx.x;
  const int ku = a.k_use;
  // The reference keeps these maxima as float (real_t); factors are 1.0/max rounded to float.
  float obj_max = (float)a.scal[0];
  float probe_max = (float)a.scal[1];
  double err = 0.0;
  const long canvas = (long)a.obj_rows * a.obj_cols;

  for (int s = 0; s < a.n_pos; ++s) {
    const int fi = a.order[s];
    const float2 pos = a.pos[fi];
    const float ascale = a.amp_scale[fi];
    const __half* amp = a.amp + (long)fi * MM;

    // psi_old_k = ROI * P_k  (bilinear ROI at the fractional position)
    for (int px = tid; px < MM; px += NT) {
      const int ii = px / M, jj = px % M;
      const float2 roi = bilinear(a.obj, a.obj_rows, a.obj_cols, (float)jj + pos.y + 0.5f, (float)ii + pos.x + 0.5f,
                                  a.tex_quantize);
      for (int k = 0; k < ku; ++k) {
        const float2 v = cmul(roi, a.probe[(long)k * MM + px]);
        a.psi_old[(long)k * MM + px] = v;
        a.psi[(long)k * MM + px] = v;
      }
    }
    __syncthreads();
    for (int k = 0; k < ku; ++k) wg_fft2<M, false, NT>(a.psi + (long)k * MM, lds, a.tw);

    // Fourier modulus constraint (d_adjustFFT / d_adjustModalFFT)
    const float nrm = 1.f / (float)MM;
    for (int px = tid; px < MM; px += NT) {
      const float A = load_amp(amp, px, ascale);
      if (ku == 1) {
        a.psi[px] = modulus_phase(a.psi[px], A, nrm);
      } else {
        float ms = 0.f;
        for (int k = 0; k < ku; ++k) ms += cabs2(a.psi[(long)k * MM + px]);
        const float f = A * rsqrtf(ms) * nrm;
        for (int k = 0; k < ku; ++k) a.psi[(long)k * MM + px] = cscale(a.psi[(long)k * MM + px], f);
      }
    }
    __syncthreads();
    for (int k = 0; k < ku; ++k) wg_fft2<M, true, NT>(a.psi + (long)k * MM, lds, a.tw);

    // Object update at the rounded position (d_updateObjectArray(M))
    const int rx = (int)roundf(pos.x), ry = (int)roundf(pos.y);
    const float co = (float)(1.0 / (double)probe_max);
    for (int px = tid; px < MM; px += NT) {
      const int ii = px / M, jj = px % M;
      float2 u = make_float2(0.f, 0.f);
      for (int k = 0; k < ku; ++k) {
        const long q = (long)k * MM + px;
        u = cadd(u, cmulcj(a.probe[q], csub(a.psi[q], a.psi_old[q])));
      }
      const long oi = (long)(rx + ii) * a.obj_cols + (ry + jj);
      float2 o = cadd(a.obj[oi], cscale(u, co));
      const float mag = sqrtf(o.x * o.x + o.y * o.y);
      if (a.phase_constraint) o = clamp_phase(o);
      a.obj[oi] = o;
      a.obj_int[oi] = mag * mag;
    }
    __syncthreads();
    {
      float mx = 0.f;
      for (long q = tid; q < canvas; q += NT) mx = fmaxf(mx, a.obj_int[q]);
      obj_max = block_max<NT>(mx, redf);
    }

    // Probe update with the updated object (d_updateProbe(M)); Probe::updateMaxIntensity
    if (a.update_probe) {
      const float cp = (float)(1.0 / (double)obj_max);
      for (int px = tid; px < MM; px += NT) {
        const int ii = px / M, jj = px % M;
        const float2 oc = a.obj[(long)(rx + ii) * a.obj_cols + (ry + jj)];
        for (int k = 0; k < ku; ++k) {
          const long q = (long)k * MM + px;
          float2 p = cadd(cscale(cmulcj(oc, csub(a.psi[q], a.psi_old[q])), cp), a.probe[q]);
          a.probe[q] = p;
          const float ia = sqrtf(p.x * p.x + p.y * p.y);
          a.probe_int[q] = ia * ia;
        }
      }
      __syncthreads();
      float mx = 0.f;
      if (a.modal) {
        for (int px = tid; px < MM; px += NT) {
          float sm = 0.f;
          for (int k = 0; k < a.k_total; ++k) sm += a.probe_int[(long)k * MM + px];
          mx = fmaxf(mx, sm);
        }
      } else {
        for (long q = tid; q < (long)a.k_total * MM; q += NT) mx = fmaxf(mx, a.probe_int[q]);
      }
      probe_max = block_max<NT>(mx, redf);
    }

    // Optional Fourier error (PhaserUtil::calculateER; ePIE.cpp:124-125)
    if (a.rms) {
      for (int k = 0; k < ku; ++k) wg_fft2<M, false, NT>(a.psi_old + (long)k * MM, lds, a.tw);
      double acc = 0.0;
      for (int px = tid; px < MM; px += NT) {
        float2 sm = make_float2(0.f, 0.f);
        for (int k = 0; k < ku; ++k) sm = cadd(sm, a.psi_old[(long)k * MM + px]);
        acc += fabs((double)load_amp(amp, px, ascale) - sqrt((double)sm.x * sm.x + (double)sm.y * sm.y));
      }
      acc = block_sum<NT>(acc, redd);
      err += acc / ku / (double)a.amp_sum[fi];
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.scal[0] = obj_max;
    a.scal[1] = probe_max;
    a.scal[2] = err;
  }
}

// threadIdx.x behind an empty asm: per-lane addresses derived from it are recomputed in each
// phase instead of being kept live (and spilled) across the transforms; the asm also bounds a
// scheduling region, so each phase calls it once.
__device__ __forceinline__ int lane() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// A f / |f| * nrm: the projection of modulus_phase without its atan2 / sincos (which would make
// the modulus step the longest of a position on one CU), branch-free: f is first scaled by an
// exact power of two so that |f|^2 cannot overflow or underflow; f = 0 gives A nrm, as there.
__device__ __forceinline__ float2 modulus_unit(float2 f, float a, float nrm) {
  const float s = fmaxf(fabsf(f.x), fabsf(f.y));
  int e;
  (void)frexpf(s, &e);
  const float xs = ldexpf(f.x, -e), ys = ldexpf(f.y, -e);
  const float k = a * nrm * rsqrtf(xs * xs + ys * ys);
  return s > 0.f ? make_float2(xs * k, ys * k) : make_float2(a * nrm, 0.f);
}

// max of the |O|^2 cache over the canvas (non-negative values) and a flat index where it is
// attained, 16 loads of 16 bytes in flight per thread
struct MaxAt {
  float v;
  long at;
};
template <int NT>
__device__ MaxAt canvas_max(const float* __restrict__ oi, long canvas, float* redf, long* redi) {
  constexpr int U = 16;
  const float4* oi4 = reinterpret_cast<const float4*>(oi);
  const long c4 = canvas >> 2;
  const int t = lane();
  float mx = 0.f;
  long ix = 0;
  for (long q0 = 0; q0 < c4; q0 += (long)U * NT) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long q = q0 + t + (long)u * NT;
      v[u] = q < c4 ? oi4[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long q = 4 * (q0 + t + (long)u * NT);
      if (v[u].x > mx) { mx = v[u].x; ix = q; }
      if (v[u].y > mx) { mx = v[u].y; ix = q + 1; }
      if (v[u].z > mx) { mx = v[u].z; ix = q + 2; }
      if (v[u].w > mx) { mx = v[u].w; ix = q + 3; }
    }
  }
  for (long q = (c4 << 2) + t; q < canvas; q += NT)
    if (oi[q] > mx) { mx = oi[q]; ix = q; }
  const float m = block_max<NT>(mx, redf);
  // a location of the maximum: the lowest index among the threads that hold it
  long cand = mx == m ? ix : canvas;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cand = min(cand, (long)__shfl_xor(cand, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) redi[threadIdx.x >> 6] = cand;
  __syncthreads();
  long best = redi[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) best = min(best, redi[i]);
  __syncthreads();
  return MaxAt{m, best};
}

// The object window [ib, ib + M + 2) x [jb, jb + M + 2) (indices clamped to the canvas) behind
// a position's bilinear taps, staged in LDS with row pitch M + 2: every tap of the frame is an
// element of it (the first tap of pixel (ii, jj) is at row ii or ii + 1 and column jj or jj + 1
// of the window when ib = floor(pos.x), jb = floor(pos.y)).
template <int M, int NT>
__device__ __forceinline__ void load_window(const EpieArgs& a, int ib, int jb, float2* win) {
  constexpr int WP = M + 2, WN = WP * WP, U = (WN + NT - 1) / NT;
  const int t = lane();
  float2 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = t + u * NT;
    const int r = min(max(ib + e / WP, 0), a.obj_rows - 1), c = min(max(jb + e % WP, 0), a.obj_cols - 1);
    v[u] = e < WN ? a.obj[(long)r * a.obj_cols + c] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = t + u * NT;
    if (e < WN) win[e] = v[u];
  }
}

// psi_old = ROI * P (bilinear ROI at the fractional position) from the staged window, into
// registers (the caller writes it into the frame after a barrier: frame and window share LDS).
template <int M, int NT, int PPT, bool QUANT>
__device__ __forceinline__ void epie_roi(const EpieArgs& a, float2 pos, int ib, int jb, const float2* win,
                                         float2 (&po)[PPT]) {
  constexpr int WP = M + 2;
  const int t = lane();
  float2 pv[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) pv[j] = a.probe[t + j * NT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int px = t + j * NT;
    const TapPos q = tap_pos<QUANT>((float)(px % M) + pos.y + 0.5f, (float)(px / M) + pos.x + 0.5f);
    const int r = min(max(q.j0 - ib, 0), M), c = min(max(q.i0 - jb, 0), M);
    const float2* w = win + r * WP + c;
    po[j] = cmul(bilinear_sum(q.al, q.be, w[0], w[1], w[WP], w[WP + 1]), pv[j]);
  }
}

// Resident-frame layout: row pitch M + 1 (conflict-free rows and columns in LDS).
template <int M>
__device__ __forceinline__ int at(int px) {
  return (px / M) * (M + 1) + (px % M);
}

// Object update at the rounded position (d_updateObjectArray) from psi - psi_old in LDS: all
// loads of G pixels before their stores (the stores could otherwise alias the next loads).
// Leaves conj(O_new) (psi - psi_old), the probe update's increment before its scale, in its
// place, and returns this thread's max of the new |O|^2 with its flat canvas index.
template <int M, int NT, int PPT, bool PC>
__device__ __forceinline__ MaxAt epie_obj_update(const EpieArgs& a, int rx, int ry, float co, float2* lds) {
  constexpr int G = 4;
  MaxAt m{0.f, 0};
#pragma unroll
  for (int j0 = 0; j0 < PPT; j0 += G) {
    const int t = lane();
    float2 ov[G], pv[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int px = t + (j0 + u) * NT;
      ov[u] = a.obj[(long)(rx + px / M) * a.obj_cols + (ry + px % M)];
      pv[u] = a.probe[px];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int px = t + (j0 + u) * NT;
      const long oi = (long)(rx + px / M) * a.obj_cols + (ry + px % M);
      const float2 d = lds[at<M>(px)];  // psi - psi_old
      float2 o = cadd(ov[u], cscale(cmulcj(pv[u], d), co));
      const float mag = sqrtf(o.x * o.x + o.y * o.y);
      if (PC) o = clamp_phase(o);
      a.obj[oi] = o;
      a.obj_int[oi] = mag * mag;
      if (mag * mag > m.v) m = MaxAt{mag * mag, oi};
      lds[at<M>(px)] = cmulcj(o, d);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  return m;
}

// Probe update with the updated object (d_updateProbe), from the increment the object update
// left in LDS; returns this thread's max |P|^2. The |P|^2 cache of the updated mode is written
// once, after the last position (nothing reads it in between).
template <int M, int NT, int PPT>
__device__ __forceinline__ float epie_probe_update(const EpieArgs& a, float cp, const float2* lds) {
  const int t = lane();
  float2 pv[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) pv[j] = a.probe[t + j * NT];
  float mx = 0.f;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int px = t + j * NT;
    const float2 p = cadd(cscale(lds[at<M>(px)], cp), pv[j]);
    a.probe[px] = p;
    const float ia = sqrtf(p.x * p.x + p.y * p.y);
    mx = fmaxf(mx, ia * ia);
  }
  return mx;
}

// Single-mode ePIE with the whole M x M exit wave resident in LDS (M (M + 1) complex values;
// the object window of the bilinear ROI, (M + 2)^2, shares the space, 132 KiB at M = 128): both
// 2D transforms and the modulus constraint run without a global round trip, and psi_old is
// never stored (the modulus step transforms the difference). Each thread owns the pixels
// px = tid + j NT of the frame in every phase. Same update equations and order as k_epie, with
// psi - psi_old = IFFT(Psi' - Psi / M^2) in place of IFFT(Psi') - psi_old. One CU moves every
// byte of a position, so the phases are arranged to move few: the ROI gathers come from the staged
// window, the probe update takes its increment from LDS, and the canvas max after an object
// update is rescanned only when the previous maximum lay inside the updated ROI (otherwise it is
// the larger of the previous maximum and the ROI's new maximum, exactly).
template <int M, int NT>
__global__ __launch_bounds__(NT) void k_epie_lds(EpieArgs a) {
  constexpr int P = M + 1;
  constexpr int FRAME = (M + 2) * (M + 2);  // >= M * P
  extern __shared__ float2 lds[];           // frame (pitch P) or window (pitch M + 2), then twiddles
  float2* tws = lds + FRAME;
  __shared__ float redf[NT / 64];
  __shared__ double redd[NT / 64];
  __shared__ long redi[NT / 64];
  constexpr int MM = M * M;
  constexpr int PPT = MM / NT;
  static_assert(MM % NT == 0, "pixels must divide among the threads");
  const int tid = threadIdx.x;
  float obj_max = (float)a.scal[0];
  float probe_max = (float)a.scal[1];
  double err = 0.0;
  const long canvas = (long)a.obj_rows * a.obj_cols;
  const float nrm = 1.f / (float)MM;
  for (int i = tid; i < PTYK_TWN; i += NT) tws[i] = a.tw[i];
  // where the current obj_max is attained (valid while the probe is updated)
  long max_at = 0;
  if (a.update_probe) {
    const MaxAt m = canvas_max<NT>(a.obj_int, canvas, redf, redi);
    obj_max = m.v;
    max_at = m.at;
  }

  for (int s = 0; s < a.n_pos; ++s) {
    const int fi = a.order[s];
    const float2 pos = a.pos[fi];
    const float ascale = a.amp_scale[fi];
    const __half* amp = a.amp + (long)fi * MM;

    const int ib = (int)floorf(pos.x), jb = (int)floorf(pos.y);
    load_window<M, NT>(a, ib, jb, lds);
    __syncthreads();
    {
      float2 po[PPT];
      if (a.tex_quantize)
        epie_roi<M, NT, PPT, true>(a, pos, ib, jb, lds, po);
      else
        epie_roi<M, NT, PPT, false>(a, pos, ib, jb, lds, po);
      __syncthreads();
      const int t = lane();
#pragma unroll
      for (int j = 0; j < PPT; ++j) lds[at<M>(t + j * NT)] = po[j];
    }
    __syncthreads();
    fft_lds<M, false, M, NT, RowLayout<M, P>, true>(lds, tws, lane());
    fft_lds<M, false, M, NT, ColLayout<P>, true>(lds, tws, lane());
    // The modulus step leaves Psi' - Psi / M^2 in place of Psi, so the inverse transform yields
    // psi - psi_old directly (the transform is linear) and psi_old need not be kept. The optional
    // Fourier error (PhaserUtil::calculateER, ePIE.cpp:124-125) is of psi_old's spectrum, which
    // is Psi itself: it is summed here instead of by a third transform.
    {
      const int t = lane();
      __half ah[PPT];
#pragma unroll
      for (int j = 0; j < PPT; ++j) ah[j] = amp[t + j * NT];
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int q = at<M>(t + j * NT);
        const float2 f = lds[q];
        const float A = __half2float(ah[j]) * ascale;
        if (a.rms) acc += fabs((double)A - sqrt((double)f.x * f.x + (double)f.y * f.y));
        lds[q] = csub(modulus_unit(f, A, nrm), cscale(f, nrm));
      }
      if (a.rms) {
        acc = block_sum<NT>(acc, redd);
        err += acc / (double)a.amp_sum[fi];
      }
    }
    __syncthreads();
    fft_lds<M, true, M, NT, RowLayout<M, P>, true>(lds, tws, lane());
    fft_lds<M, true, M, NT, ColLayout<P>, true>(lds, tws, lane());

    const int rx = (int)roundf(pos.x), ry = (int)roundf(pos.y);
    const float co = (float)(1.0 / (double)probe_max);
    const MaxAt mo = a.phase_constraint ? epie_obj_update<M, NT, PPT, true>(a, rx, ry, co, lds)
                                        : epie_obj_update<M, NT, PPT, false>(a, rx, ry, co, lds);
    __syncthreads();

    // the probe update needs max |O|^2 after this position, which otherwise is only reported
    // after the last one
    if (a.update_probe) {
      const long mr = max_at / a.obj_cols, mc = max_at % a.obj_cols;
      if (mr >= rx && mr < rx + M && mc >= ry && mc < ry + M) {
        const MaxAt m = canvas_max<NT>(a.obj_int, canvas, redf, redi);
        obj_max = m.v;
        max_at = m.at;
      } else {
        const float m = block_max<NT>(mo.v, redf);
        if (m > obj_max) {
          // a location of the new maximum: the lowest index among the threads that hold it
          long cand = mo.v == m ? mo.at : canvas;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) cand = min(cand, (long)__shfl_xor(cand, o, 64));
          if ((tid & 63) == 0) redi[tid >> 6] = cand;
          __syncthreads();
          long best = redi[0];
#pragma unroll
          for (int i = 1; i < NT / 64; ++i) best = min(best, redi[i]);
          __syncthreads();
          obj_max = m;
          max_at = best;
        }
      }
      const float cp = (float)(1.0 / (double)obj_max);
      float mx = epie_probe_update<M, NT, PPT>(a, cp, lds);
      // the other (unused) modes keep their intensities
      for (long q = MM + lane(); q < (long)a.k_total * MM; q += NT) mx = fmaxf(mx, a.probe_int[q]);
      probe_max = block_max<NT>(mx, redf);
    }

    __syncthreads();
  }
  if (!a.update_probe && a.n_pos > 0) obj_max = canvas_max<NT>(a.obj_int, canvas, redf, redi).v;
  if (a.update_probe && a.n_pos > 0) {
    // each thread rereads only the probe pixels it wrote itself
    const int t = lane();
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const float2 p = a.probe[t + j * NT];
      const float ia = sqrtf(p.x * p.x + p.y * p.y);
      a.probe_int[t + j * NT] = ia * ia;
    }
  }
  if (tid == 0) {
    a.scal[0] = obj_max;
    a.scal[1] = probe_max;
    a.scal[2] = err;
  }
}

template <int M>
static hipError_t epie_impl(const EpieArgs& a, hipStream_t s) {
  const size_t lds = (size_t)WgFft<M, PTYK_EPIE_NT>::LDS * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void*)k_epie<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_epie<M>), dim3(1), dim3(PTYK_EPIE_NT), lds, s, a);
  return hipGetLastError();
}

// Frames whose exit wave fits in LDS take the resident form (single mode only).
template <int M, int NT>
static hipError_t epie_lds_impl(const EpieArgs& a, hipStream_t s) {
  const size_t lds = (size_t)((M + 2) * (M + 2) + PTYK_TWN) * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void*)k_epie_lds<M, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_epie_lds<M, NT>), dim3(1), dim3(NT), lds, s, a);
  return hipGetLastError();
}

hipError_t epie_iteration(const EpieArgs& a, hipStream_t s) {
  if (a.k_use == 1 && a.m == 128) return epie_lds_impl<128, 1024>(a, s);
  if (a.k_use == 1 && a.m == 64) return epie_lds_impl<64, 512>(a, s);
  switch (a.m) {
    case 16: return epie_impl<16>(a, s);
    case 32: return epie_impl<32>(a, s);
    case 64: return epie_impl<64>(a, s);
    case 128: return epie_impl<128>(a, s);
    case 256: return epie_impl<256>(a, s);
    case 512: return epie_impl<512>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ptyk
