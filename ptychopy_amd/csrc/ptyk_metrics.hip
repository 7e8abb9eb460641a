// Normalised object error against a known object (no reference counterpart: the reference's only
// real-space metric, d_realSpaceER (utilitiesKernels.cu:714-735, 1927-1941), compares magnitudes
// and is never called).
//
//   e(O) = min_gamma || gamma O - T ||_2 / || T ||_2,  gamma = <O, T> / <O, O>,
//
// over the illuminated region: pixels where the scan's total probe intensity
// sum_j sum_k |P_k|^2 (each frame at its ROI) exceeds a fraction of its maximum.  All sums are
// FP64 in a fixed order (block partials, then one block).
#include <hip/hip_runtime.h>

#include <vector>

#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

namespace {

constexpr int NT = 256;

// Coverage map: sum over frames covering the pixel of sum_k |P_k|^2 at the local offset.
__global__ __launch_bounds__(NT) void k_coverage(float* __restrict__ cov, const float* __restrict__ pint,
                                                 const int2* __restrict__ roi, int n, int m, int rows, int cols) {
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
      const int lrw = r - lr[j].x, lc = c - lr[j].y;
      if (lrw >= 0 && lrw < m && lc >= 0 && lc < m) s += pint[lrw * m + lc];
    }
  }
  if (r < rows && c < cols) cov[(long)r * cols + c] = s;
}

// sum_k |P_k|^2 per probe pixel
__global__ void k_mode_power(const float2* __restrict__ probe, int k, int mm, float* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  float s = 0.f;
  for (int i = 0; i < k; ++i) s += cabs2(probe[(long)i * mm + q]);
  out[q] = s;
}

__global__ __launch_bounds__(NT) void k_cov_max(const float* __restrict__ cov, long count, double* __restrict__ part) {
  __shared__ float red[NT / 64];
  float mx = 0.f;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < count; i += (long)gridDim.x * NT) mx = fmaxf(mx, cov[i]);
  mx = block_max<NT>(mx, red);
  if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

// pass 1: [sum |O|^2, sum |T|^2, Re <O,T>, Im <O,T>, count]; pass 2: sum |gamma O - T|^2
__global__ __launch_bounds__(NT) void k_err_sums(const float2* __restrict__ o, const float2* __restrict__ t,
                                                 const float* __restrict__ cov, long count, float thr, int pass,
                                                 double gr, double gi, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  double s[5] = {0, 0, 0, 0, 0};
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < count; i += (long)gridDim.x * NT) {
    if (!(cov[i] > thr)) continue;
    const double ox = o[i].x, oy = o[i].y, tx = t[i].x, ty = t[i].y;
    if (pass == 1) {
      s[0] += ox * ox + oy * oy;
      s[1] += tx * tx + ty * ty;
      s[2] += ox * tx + oy * ty;  // conj(O) T
      s[3] += ox * ty - oy * tx;
      s[4] += 1.0;
    } else {
      const double dx = gr * ox - gi * oy - tx, dy = gr * oy + gi * ox - ty;
      s[0] += dx * dx + dy * dy;
    }
  }
  const int nv = pass == 1 ? 5 : 1;
  for (int v = 0; v < nv; ++v) {
    const double r = block_sum<NT>(s[v], red);
    if (threadIdx.x == 0) part[(long)blockIdx.x * 5 + v] = r;
  }
}

// Weighted patch average: O[p] = sum_j w(p - r_j) a_j e^{i phi_j} / sum_j w(p - r_j) over the
// selected frames j covering p (ascending selection order; frame sel[j], ROI origin
// roi[sel[j]]), w = sum_k |P_k|^2; pixels whose weight sum is at most thr keep `fill`.
// Patches: n x 2 x M x M (amplitude, phase), patch j for selected frame j.
__global__ __launch_bounds__(NT) void k_stitch(float2* __restrict__ obj, const float* __restrict__ patches,
                                               const float* __restrict__ pint, const int2* __restrict__ roi,
                                               const int* __restrict__ sel, int n, int m, int rows, int cols,
                                               float thr, float2 fill) {
  __shared__ int lj[NT];
  __shared__ int wcnt[NT / 64];
  const int r0 = blockIdx.y * 4, c0 = blockIdx.x * 64;
  const int r = r0 + (threadIdx.x >> 6), c = c0 + (threadIdx.x & 63);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long mm = (long)m * m;
  float nx = 0.f, ny = 0.f, den = 0.f;
  for (int base = 0; base < n; base += NT) {
    const int f = base + threadIdx.x;
    bool hit = false;
    if (f < n) {
      const int2 o = roi[sel[f]];
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
    if (hit) lj[off + __popcll(bal & ((1ull << lane) - 1ull))] = f;
    __syncthreads();
    for (int t = 0; t < tot; ++t) {
      const int j = lj[t];
      const int2 o = roi[sel[j]];
      const int lr = r - o.x, lc = c - o.y;
      if (lr >= 0 && lr < m && lc >= 0 && lc < m) {
        const long q = (long)lr * m + lc;
        const float wt = pint[q];
        const float a = patches[(2 * (long)j) * mm + q], ph = patches[(2 * (long)j + 1) * mm + q];
        float sn, cs;
        sincosf(ph, &sn, &cs);
        nx += wt * a * cs;
        ny += wt * a * sn;
        den += wt;
      }
    }
    __syncthreads();
  }
  if (r < rows && c < cols) obj[(long)r * cols + c] = den > thr ? make_float2(nx / den, ny / den) : fill;
}

}  // namespace

hipError_t stitch_patches(float2* obj, const float* patches, const float2* probe, int k, int m, const int2* roi,
                          const int* sel, int n, int rows, int cols, float* pint, float rel_thr, float2 fill,
                          hipStream_t s) {
  const int mm = m * m;
  hipLaunchKernelGGL(k_mode_power, dim3((mm + NT - 1) / NT), dim3(NT), 0, s, probe, k, mm, pint);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::vector<float> h(mm);
  if ((e = hipMemcpyAsync(h.data(), pint, mm * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  float mx = 0.f;
  for (float v : h) mx = std::max(mx, v);
  hipLaunchKernelGGL(k_stitch, dim3((cols + 63) / 64, (rows + 3) / 4), dim3(NT), 0, s, obj, patches, pint, roi, sel,
                     n, m, rows, cols, rel_thr * mx, fill);
  return hipGetLastError();
}

hipError_t object_error(const float2* obj, const float2* truth, const float2* probe, int k, int m, const int2* roi,
                        int n, int rows, int cols, double frac, float* cov, float* pint, double* part, int part_cap,
                        double* out, hipStream_t s) {
  const long canvas = (long)rows * cols;
  const int mm = m * m;
  hipLaunchKernelGGL(k_mode_power, dim3((mm + NT - 1) / NT), dim3(NT), 0, s, probe, k, mm, pint);
  hipLaunchKernelGGL(k_coverage, dim3((cols + 63) / 64, (rows + 3) / 4), dim3(NT), 0, s, cov, pint, roi, n, m, rows,
                     cols);
  const int nb = (int)std::min<long>(std::min<long>((canvas + NT - 1) / NT, 1024), part_cap / 5);
  hipLaunchKernelGGL(k_cov_max, dim3(nb), dim3(NT), 0, s, cov, canvas, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::vector<double> h((size_t)nb * 5);
  if ((e = hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  double mx = 0.0;
  for (int i = 0; i < nb; ++i) mx = std::max(mx, h[i]);
  const float thr = (float)(frac * mx);
  hipLaunchKernelGGL(k_err_sums, dim3(nb), dim3(NT), 0, s, obj, truth, cov, canvas, thr, 1, 0.0, 0.0, part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(h.data(), part, (size_t)nb * 5 * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  double so = 0, st = 0, cr = 0, ci = 0, cnt = 0;
  for (int i = 0; i < nb; ++i) {
    so += h[5 * i];
    st += h[5 * i + 1];
    cr += h[5 * i + 2];
    ci += h[5 * i + 3];
    cnt += h[5 * i + 4];
  }
  const double gr = so > 0 ? cr / so : 0.0, gi = so > 0 ? ci / so : 0.0;
  hipLaunchKernelGGL(k_err_sums, dim3(nb), dim3(NT), 0, s, obj, truth, cov, canvas, thr, 2, gr, gi, part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(h.data(), part, (size_t)nb * 5 * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  double sd = 0;
  for (int i = 0; i < nb; ++i) sd += h[5 * i];
  out[0] = st > 0 ? std::sqrt(sd / st) : 0.0;
  out[1] = gr;
  out[2] = gi;
  out[3] = cnt;
  return hipSuccess;
}

}  // namespace ptyk
