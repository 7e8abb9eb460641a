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

int ptycho_mesh_cartesian(const ptycho_mesh_params* p, float* pos_out, int32_t* canvas_out, float* extent_out) {
  if (!p || !pos_out || p->nx <= 0 || p->ny <= 0 || p->probe_size <= 0 || !(p->dx_s > 0)) return -1;
  const int nx = p->nx, ny = p->ny;
  const float step_x = (float)p->step_x, step_y = (float)p->step_y;  // ExperimentParams::stepSize is float2
  const float one_over_dxs = (float)(1.0 / p->dx_s);                  // real_t argument of d_generateMesh
  const float drift_x = (float)p->drift_x, drift_y = (float)p->drift_y;
  const int jr = p->jitter_radius;
  std::mt19937_64 gen(p->seed);
  std::vector<float> px((size_t)nx * ny), py((size_t)nx * ny);
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
      px[(size_t)i * ny + j] = std::fmaf(vx, one_over_dxs, jx);
      py[(size_t)i * ny + j] = std::fmaf(vy, one_over_dxs, jy);
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
