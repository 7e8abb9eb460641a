// Scan geometry on the host, with the reference's float32 arithmetic reproduced bit for bit.
//
//  * sample-plane pixel size dx_s = lambda z / (size dx_d), rounded to float
//    (Parameters::samplePlanePixelSize, Parameters.h:175-179; stored back as double,
//    Parameters.cpp:650-652);
//  * Cartesian positions (d_generateMesh<1>, ScanMeshKernels.cu:46-93, launched by
//    h_generateCartesianMesh, ScanMeshKernels.cu:154-164, for CartesianScanMesh::generateMesh,
//    CartesianScanMesh.cpp:57-92);
//  * extents over the pitched position buffer, whose row padding holds zeros
//    (h_maxFloat2 / h_minFloat2 with FLT_MIN / FLT_MAX identities, utilitiesKernels.cu:1291-1301;
//    rows padded to a multiple of 32, Cuda2DArray.hpp:67-74), the canvas
//    ceil(max - min + M) (IPtychoScanMesh::calculateMeshDimensions, IPtychoScanMesh.cpp:86-99)
//    and the shift pos - min with optional axis flip (addScanPosition, IPtychoScanMesh.cpp:102-109).
// Jitter: the reference draws curand_uniform (time-seeded); here a seeded mt19937_64 supplies
// u in (0, 1] and the jitter keeps the reference's form r - u * 2r per axis.
// nvcc contracts a*b + c into one fused multiply-add by default, so the device arithmetic of
// d_generateMesh is reproduced with explicit fmaf; this file is compiled with contraction off.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "ptycho.h"

namespace {
inline float uniform_open0(std::mt19937_64& g) {
  // (0, 1]: 53 random bits, shifted by one ulp so 0 is excluded and 1 included
  return (float)(((double)(g() >> 11) + 1.0) * (1.0 / 9007199254740992.0));
}
}  // namespace

extern "C" {

double ptycho_sample_pixel_size(double lambda, double z, int32_t size, double dx_d) {
  const float v = (float)((lambda * z) / ((double)size * dx_d));
  return (double)v;
}

// Unshifted Cartesian positions in pixels, row-major nx x ny (d_generateMesh<1> arithmetic).
int ptycho_mesh_cartesian_raw(const ptycho_mesh_params* p, float* raw_out) {
  if (!p || !raw_out || p->nx <= 0 || p->ny <= 0 || p->probe_size <= 0 || !(p->dx_s > 0)) return -1;
  const int nx = p->nx, ny = p->ny;
  const float step_x = (float)p->step_x, step_y = (float)p->step_y;  // ExperimentParams::stepSize is float2
  const float one_over_dxs = (float)(1.0 / p->dx_s);                  // real_t argument of d_generateMesh
  const float drift_x = (float)p->drift_x, drift_y = (float)p->drift_y;
  const int jr = p->jitter_radius;
  std::mt19937_64 gen(p->seed);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      float jx = 0.f, jy = 0.f;
      if (jr > 0) {
        jx = std::fmaf(-(uniform_open0(gen) * 2.0f), (float)jr, (float)jr);
        jy = std::fmaf(-(uniform_open0(gen) * 2.0f), (float)jr, (float)jr);
      }
      float m = (float)i, d = (float)(nx - 1) * 0.5f;
      float vx = (p->mirror_x ? (d - m) : (m - d)) * step_x;
      m = (float)j;
      d = (float)(ny - 1) * 0.5f;
      float vy = (p->mirror_y ? (d - m) : (m - d)) * step_y;
      vx = std::fmaf((float)i, drift_x, vx);
      vy = std::fmaf((float)j, drift_y, vy);
      raw_out[2 * ((size_t)i * ny + j)] = std::fmaf(vx, one_over_dxs, jx);
      raw_out[2 * ((size_t)i * ny + j) + 1] = std::fmaf(vy, one_over_dxs, jy);
    }
  return 0;
}

int ptycho_mesh_cartesian(const ptycho_mesh_params* p, float* pos_out, int32_t* canvas_out, float* extent_out) {
  if (!p || !pos_out || p->nx <= 0 || p->ny <= 0 || p->probe_size <= 0 || !(p->dx_s > 0)) return -1;
  const int nx = p->nx, ny = p->ny;
  std::vector<float> raw(2 * (size_t)nx * ny);
  if (ptycho_mesh_cartesian_raw(p, raw.data()) != 0) return -1;
  std::vector<float> px((size_t)nx * ny), py((size_t)nx * ny);
  for (size_t q = 0; q < px.size(); ++q) {
    px[q] = raw[2 * q];
    py[q] = raw[2 * q + 1];
  }
  // extents over the pitched buffer: rows of ny values padded with zeros to a multiple of 32
  float mxx = FLT_MIN, mxy = FLT_MIN, mnx = FLT_MAX, mny = FLT_MAX;
  for (size_t q = 0; q < px.size(); ++q) {
    mxx = std::max(mxx, px[q]);
    mxy = std::max(mxy, py[q]);
    mnx = std::min(mnx, px[q]);
    mny = std::min(mny, py[q]);
  }
  if (ny % 32 != 0) {
    mxx = std::max(mxx, 0.f);
    mxy = std::max(mxy, 0.f);
    mnx = std::min(mnx, 0.f);
    mny = std::min(mny, 0.f);
  }
  const int32_t cx = (int32_t)std::ceil((mxx - mnx) + (float)p->probe_size);
  const int32_t cy = (int32_t)std::ceil((mxy - mny) + (float)p->probe_size);
  for (size_t q = 0; q < px.size(); ++q) {
    const float x = px[q], y = py[q];
    pos_out[2 * q] = p->flip ? y - mny : x - mnx;
    pos_out[2 * q + 1] = p->flip ? x - mnx : y - mny;
  }
  if (canvas_out) {
    canvas_out[0] = cx;
    canvas_out[1] = cy;
  }
  if (extent_out) {
    extent_out[0] = mnx;
    extent_out[1] = mny;
    extent_out[2] = mxx;
    extent_out[3] = mxy;
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- MLs geometry
extern "C" int ptycho_mls_canvas(const float* pos, int64_t n, const float* extent, int32_t probe_size,
                                 int32_t* canvas_out, int32_t* box_out) {
  if (!pos || !extent || n <= 0 || probe_size <= 0) return -1;
  // generateMeshML (CartesianScanMesh.cpp:118-172): ROI bounds of every position on a 1e5 x 1e5
  // sentinel canvas, centred on the middle of the extents.  The reference takes the row extent
  // from the column coordinates and vice versa (pos.y paired with ceilvar.x), which leaves the
  // ROIs of a non-square scan outside its canvas; here rows come from the row coordinates (the
  // same canvas for square scans).
  const float Mf = (float)probe_size;
  const float cv = (float)std::ceil(1e5f * 1.0 / 2 - probe_size * 1.0 / 2);
  const float mnx = extent[0], mny = extent[1], mxx = extent[2], mxy = extent[3];
  const float rx = (mxx + mnx) / 2, ry = (mxy + mny) / 2;
  unsigned lox = (unsigned)1e5f, loy = (unsigned)1e5f, hix = 0, hiy = 0;
  for (int64_t q = 0; q < n; ++q) {
    const unsigned tx = (unsigned)std::round(pos[2 * q] - rx + cv);
    const unsigned ty = (unsigned)std::round(pos[2 * q + 1] - ry + cv);
    lox = std::min(lox, tx - 1);
    hix = std::max(hix, (unsigned)((float)tx + Mf - 2));
    loy = std::min(loy, ty - 1);
    hiy = std::max(hiy, (unsigned)((float)ty + Mf - 2));
  }
  const float extra = 1 + 0.2f;
  const unsigned npx = (unsigned)std::ceil((float)(hix - lox) * extra);
  const unsigned npy = (unsigned)std::ceil((float)(hiy - loy) * extra);
  if (canvas_out) {
    canvas_out[0] = (int32_t)npx;
    canvas_out[1] = (int32_t)npy;
  }
  // IPtychoScanMesh::precalculate_ROI (IPtychoScanMesh.cpp:179-184), 1-based inclusive
  if (box_out) {
    box_out[0] = (int32_t)std::round(npx * 0.5 + (double)mnx);
    box_out[1] = (int32_t)std::round(npx * 0.5 + (double)mxx) - 1;
    box_out[2] = (int32_t)std::round(npy * 0.5 + (double)mny);
    box_out[3] = (int32_t)std::round(npy * 0.5 + (double)mxy) - 1;
  }
  return 0;
}

// IPtychoScanMesh::get_nonoverlapping_indices (IPtychoScanMesh.cpp:826-913): min(Niter, 30)
// random partitions of the frames into groups of floor(N/100) (at least 1), each partition's
// groups ordered by decreasing size (ties in creation order).  random_shuffle becomes a seeded
// Fisher-Yates shuffle.  Null outputs: only the counts.
extern "C" int ptycho_mls_partitions(int64_t n, int32_t niter, uint64_t seed, int32_t* frames_out,
                                     int32_t* offsets_out, int32_t* n_groups, int32_t* n_partitions) {
  if (n <= 0 || niter <= 0 || !n_groups || !n_partitions) return -1;
  const int64_t grouping = std::max<int64_t>(1, n / 100);
  const int32_t groups = (int32_t)((n + grouping - 1) / grouping);
  const int32_t parts = std::min(niter, 30);
  *n_groups = groups;
  *n_partitions = parts;
  if (!frames_out || !offsets_out) return 0;
  std::mt19937_64 g(seed);
  std::vector<int32_t> idx(n);
  for (int32_t p = 0; p < parts; ++p) {
    for (int64_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
    for (int64_t i = n - 1; i > 0; --i) std::swap(idx[i], idx[g() % (uint64_t)(i + 1)]);
    // groups j cover idx[j*grouping, min(n, (j+1)*grouping)); sizes are non-increasing in j
    int32_t* fr = frames_out + (size_t)p * n;
    int32_t* of = offsets_out + (size_t)p * (groups + 1);
    of[0] = 0;
    for (int32_t j = 0; j < groups; ++j) {
      const int64_t b = j * grouping, e = std::min<int64_t>(n, (j + 1) * grouping);
      for (int64_t i = b; i < e; ++i) fr[i] = idx[i];
      of[j + 1] = (int32_t)e;
    }
  }
  return 0;
}

// IPtychoScanMesh::get_close_indices (IPtychoScanMesh.cpp:422-824; the "MLC" alternative to
// get_nonoverlapping_indices that IPhaser.cpp:159 leaves commented out): spatially compact groups
// of about floor(N/100) frames.  The reference's body reads its starting groups and distances from
// fixed files on its authors' disk; this follows the steps it takes around them:
//  1. k-means of the positions into ceil(N/grouping) clusters: k-means++ seeding (utilities.h:225-251)
//     and Lloyd iterations until at most N/1024 points change (utilities.h:253-285), ten seeded
//     restarts, keeping the one whose cluster sizes have the smallest standard deviation
//     (scanPositionSD, IPtychoScanMesh.cpp:320-333);
//  2. balancing: while a cluster holds more than `grouping` frames, the frame of an over-full cluster
//     closest to the centre of the smallest cluster moves there (IPtychoScanMesh.cpp:564-628), until
//     at most two sizes remain;
//  3. centres become the per-cluster medians of x and y (CalcMHWScore, IPtychoScanMesh.cpp:399-420)
//     and, for up to ten passes while the share of frames outside their nearest cluster falls, the
//     frame farthest from its centre swaps clusters with the frame of its nearest cluster closest to
//     the first one's centre (sizes are kept; IPtychoScanMesh.cpp:656-754);
//  4. groups in order of decreasing size (ties by cluster index), frames ascending
//     (IPtychoScanMesh.cpp:756-822).
// pos: n x 2 (any unit).  Null outputs: only the count.
extern "C" int ptycho_mls_compact_groups(const float* pos, int64_t n, uint64_t seed, int32_t* frames_out,
                                         int32_t* offsets_out, int32_t* n_groups) {
  if (!pos || n <= 0 || !n_groups) return -1;
  const int64_t grouping = std::max<int64_t>(1, n / 100);
  const int K = (int)((n + grouping - 1) / grouping);
  *n_groups = K;
  if (!frames_out || !offsets_out) return 0;
  auto d2 = [&](int64_t i, double cx, double cy) {
    const double dx = pos[2 * i] - cx, dy = pos[2 * i + 1] - cy;
    return dx * dx + dy * dy;
  };
  std::mt19937_64 g(seed);
  std::vector<int> best_group;
  std::vector<double> best_cx, best_cy;
  double best_sd = HUGE_VAL;
  std::vector<int> grp(n);
  std::vector<double> cx(K), cy(K), dmin(n);
  for (int rep = 0; rep < 10; ++rep) {
    // k-means++ seeding
    int64_t first = (int64_t)(g() % (uint64_t)n);
    cx[0] = pos[2 * first];
    cy[0] = pos[2 * first + 1];
    for (int64_t i = 0; i < n; ++i) dmin[i] = d2(i, cx[0], cy[0]);
    for (int k = 1; k < K; ++k) {
      double sum = 0.0;
      for (int64_t i = 0; i < n; ++i) sum += dmin[i];
      double r = (double)(g() >> 11) * (1.0 / 9007199254740992.0) * sum;
      int64_t pick = n - 1;
      for (int64_t i = 0; i < n; ++i) {
        r -= dmin[i];
        if (r <= 0) {
          pick = i;
          break;
        }
      }
      cx[k] = pos[2 * pick];
      cy[k] = pos[2 * pick + 1];
      for (int64_t i = 0; i < n; ++i) dmin[i] = std::min(dmin[i], d2(i, cx[k], cy[k]));
    }
    auto nearest = [&](int64_t i) {
      int bi = 0;
      double bd = HUGE_VAL;
      for (int k = 0; k < K; ++k) {
        const double d = d2(i, cx[k], cy[k]);
        if (d < bd) {
          bd = d;
          bi = k;
        }
      }
      return bi;
    };
    for (int64_t i = 0; i < n; ++i) grp[i] = nearest(i);
    // Lloyd iterations
    for (int it = 0; it < 1000; ++it) {
      std::vector<double> sx(K, 0.0), sy(K, 0.0);
      std::vector<int64_t> cnt(K, 0);
      for (int64_t i = 0; i < n; ++i) {
        sx[grp[i]] += pos[2 * i];
        sy[grp[i]] += pos[2 * i + 1];
        cnt[grp[i]]++;
      }
      for (int k = 0; k < K; ++k)
        if (cnt[k]) {
          cx[k] = sx[k] / cnt[k];
          cy[k] = sy[k] / cnt[k];
        }
      int64_t changed = 0;
      for (int64_t i = 0; i < n; ++i) {
        const int b = nearest(i);
        if (b != grp[i]) {
          grp[i] = b;
          ++changed;
        }
      }
      if (changed <= (n >> 10)) break;
    }
    std::vector<int64_t> cnt(K, 0);
    for (int64_t i = 0; i < n; ++i) cnt[grp[i]]++;
    double mean = (double)n / K, var = 0.0;
    for (int k = 0; k < K; ++k) var += ((double)cnt[k] - mean) * ((double)cnt[k] - mean);
    const double sd = std::sqrt(var / K);
    if (sd < best_sd) {
      best_sd = sd;
      best_group = grp;
      best_cx = cx;
      best_cy = cy;
    }
  }
  grp = best_group;
  cx = best_cx;
  cy = best_cy;
  std::vector<int64_t> cnt(K, 0);
  for (int64_t i = 0; i < n; ++i) cnt[grp[i]]++;
  // balancing
  for (int64_t it = 1; it <= 4 * n + 1000; ++it) {
    std::vector<int64_t> sizes(cnt);
    std::sort(sizes.begin(), sizes.end());
    const int64_t distinct = std::unique(sizes.begin(), sizes.end()) - sizes.begin();
    if (distinct <= 2 && ((int64_t)K * grouping != n || it > 1000)) break;
    if (distinct == 1) break;
    const int smallest = (int)(std::min_element(cnt.begin(), cnt.end()) - cnt.begin());
    int64_t pick = -1;
    double pd = HUGE_VAL;
    for (int64_t i = 0; i < n; ++i)
      if (cnt[grp[i]] > grouping) {
        const double d = d2(i, cx[smallest], cy[smallest]);
        if (d < pd) {
          pd = d;
          pick = i;
        }
      }
    if (pick < 0) break;
    cnt[grp[pick]]--;
    grp[pick] = smallest;
    cnt[smallest]++;
  }
  // median centres, then size-preserving swaps toward the nearest centre
  for (int k = 0; k < K; ++k) {
    std::vector<double> xs, ys;
    for (int64_t i = 0; i < n; ++i)
      if (grp[i] == k) {
        xs.push_back(pos[2 * i]);
        ys.push_back(pos[2 * i + 1]);
      }
    if (xs.empty()) continue;
    std::sort(xs.begin(), xs.end());
    std::sort(ys.begin(), ys.end());
    const size_t s = xs.size();
    cx[k] = s % 2 ? xs[s / 2] : 0.5 * (xs[s / 2 - 1] + xs[s / 2]);
    cy[k] = s % 2 ? ys[s / 2] : 0.5 * (ys[s / 2 - 1] + ys[s / 2]);
  }
  std::vector<int> opt(n);
  for (int64_t i = 0; i < n; ++i) {
    int bi = 0;
    double bd = HUGE_VAL;
    for (int k = 0; k < K; ++k) {
      const double d = d2(i, cx[k], cy[k]);
      if (d < bd) {
        bd = d;
        bi = k;
      }
    }
    opt[i] = bi;
  }
  double ratio0 = 1.0;
  for (int pass = 0; pass < 10; ++pass) {
    std::vector<int64_t> sw;
    for (int64_t i = 0; i < n; ++i)
      if (grp[i] != opt[i]) sw.push_back(i);
    const double ratio = (double)sw.size() / n;
    if (ratio >= ratio0) break;
    ratio0 = ratio;
    std::vector<char> done(sw.size(), 0);
    for (size_t step = 0; step < sw.size(); ++step) {
      int64_t worst = -1;
      size_t wi = 0;
      double wd = -1.0;
      for (size_t s = 0; s < sw.size(); ++s)
        if (!done[s]) {
          const double d = d2(sw[s], cx[grp[sw[s]]], cy[grp[sw[s]]]);
          if (d > wd) {
            wd = d;
            worst = sw[s];
            wi = s;
          }
        }
      if (worst < 0) break;
      done[wi] = 1;
      const int g_old = grp[worst], g_new = opt[worst];
      if (g_old == g_new) continue;
      int64_t partner = -1;
      double bd = HUGE_VAL;
      for (int64_t i = 0; i < n; ++i)
        if (grp[i] == g_new) {
          const double d = d2(i, cx[g_old], cy[g_old]);
          if (d < bd) {
            bd = d;
            partner = i;
          }
        }
      if (partner < 0) continue;
      grp[worst] = g_new;
      grp[partner] = g_old;
      for (size_t s = 0; s < sw.size(); ++s)
        if (sw[s] == partner) done[s] = 1;
    }
  }
  // output: groups by decreasing size (ties by index), frames ascending
  std::fill(cnt.begin(), cnt.end(), 0);
  for (int64_t i = 0; i < n; ++i) cnt[grp[i]]++;
  std::vector<int> order(K);
  for (int k = 0; k < K; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
  int64_t off = 0;
  offsets_out[0] = 0;
  for (int o = 0; o < K; ++o) {
    for (int64_t i = 0; i < n; ++i)
      if (grp[i] == order[o]) frames_out[off++] = (int32_t)i;
    offsets_out[o + 1] = (int32_t)off;
  }
  return 0;
}
