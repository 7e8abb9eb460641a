// Host-side driver implementation (see phaser.h).  Compiled as C++ with contraction off; the
// mesh arithmetic reproduces nvcc's fused multiply-adds explicitly.
#include "phaser.h"

#include <sys/stat.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "ptycho.h"

namespace ptycho {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline float uniform_open0(std::mt19937_64& g) {
  return (float)(((double)(g() >> 11) + 1.0) * (1.0 / 9007199254740992.0));
}
bool parse_pair(const std::string& v, double* a, double* b, bool* pair) {
  const size_t c = v.find(',');
  *pair = c != std::string::npos;
  *a = atof(v.substr(0, c).c_str());
  *b = *pair ? atof(v.substr(c + 1).c_str()) : *a;
  return true;
}
}  // namespace

// ---------------------------------------------------------------- parameters
void Parameters::finalize() {
  p.symmetric_array_size = r.desiredShape;
  e.dx_s = ptycho_sample_pixel_size(e.lambda, e.z_d, (int32_t)r.desiredShape, e.dx_d);
  e.dx = e.dx_s;
}

bool Parameters::parseCommandLine(const std::vector<std::string>& args, std::string* err) {
  for (const std::string& tok : args) {
    std::string t = tok;
    while (!t.empty() && t[0] == '-') t.erase(0, 1);
    if (t.empty()) continue;
    const size_t eq = t.find('=');
    const std::string key = t.substr(0, eq), val = eq == std::string::npos ? "" : t.substr(eq + 1);
    const int iv = atoi(val.c_str());
    double a = 0, b = 0;
    bool pair = false;
    if (key == "sqrtData") p.flags |= 1;
    else if (key == "fftShiftData") p.flags |= 4;
    else if (key == "beamstopMask") p.beamstopMask = iv > 0;
    else if (key == "blind") r.blind = iv > 0;
    else if (key == "simulate") r.simulated = iv > 0;
    else if (key == "spiralScan") e.spiralScan = iv > 0;
    else if (key == "bitDepth") e.bit_depth = iv;
    else if (key == "flipScanAxis") r.flipScanAxis = iv > 0;
    else if (key == "mirror1stScanAxis") r.mirrorX = iv > 0;
    else if (key == "mirror2ndScanAxis") r.mirrorY = iv > 0;
    else if (key == "RMS") r.calculateRMS = iv > 0;
    else if (key == "binaryOutput") r.binaryOutput = iv > 0;
    else if (key == "threads") p.io_threads = iv;
    else if (key == "i") r.iterations = iv;
    else if (key == "save") r.save = iv;
    else if (key == "T") r.time = atof(val.c_str());
    else if (key == "size") r.desiredShape = iv;
    else if (key == "PPS") r.probe_pos_search = iv;
    else if (key == "probeModes") r.probeModes = r.nProbes = iv;
    else if (key == "fs") p.fileStartIndex = iv;
    else if (key == "dpf") p.diffsPerFile = iv;
    else if (key == "angles") p.projectionsNum = iv;
    else if (key == "threshold") p.threshold_raw_data = (float)iv;
    else if (key == "jitterRadius") r.jitterRadius = iv;
    else if (key == "rotate90") p.rotate_90_times = iv;
    else if (key == "overlap") r.halo = iv;
    else if (key == "shareFrequency") r.shareFrequency = iv;
    else if (key == "updateProbe") r.updateProbe = iv;
    else if (key == "updateModes") r.updateProbeModes = iv;
    else if (key == "updateVis") r.updateVis = iv;
    else if (key == "phaseConstraint") r.phaseConstraint = iv;
    else if (key == "fp") p.dataFilePattern = val;
    else if (key == "hdf5path") p.hdf5DatasetName = val;
    else if (key == "scanDims") {
      parse_pair(val, &a, &b, &pair);
      e.scanDims_x = (unsigned)a;
      e.scanDims_y = pair ? (unsigned)b : 1;
    } else if (key == "drift") {
      parse_pair(val, &a, &b, &pair);
      e.drift_x = (float)a;
      e.drift_y = (float)b;
    } else if (key == "qxy") {
      parse_pair(val, &a, &b, &pair);
      p.symmetric_array_center_x = (int)a;
      p.symmetric_array_center_y = (int)b;
    } else if (key == "nxy") {
      parse_pair(val, &a, &b, &pair);
      p.raw_file_size_x = (int)a;
      p.raw_file_size_y = (int)b;
    } else if (key == "step") {
      parse_pair(val, &a, &b, &pair);
      e.stepSize_x = (float)a;
      e.stepSize_y = (float)b;
    } else if (key == "beamSize") e.beamSize = atof(val.c_str());
    else if (key == "lambda") e.lambda = atof(val.c_str());
    else if (key == "dx_d") e.dx_d = atof(val.c_str());
    else if (key == "z") e.z_d = atof(val.c_str());
    else if (key == "jobID") r.reconstructionID = val;
    else if (key == "algorithm") r.algorithm = val;
    else if (key == "lf") r.positionsFilename = val;
    else if (key == "probeGuess") r.probeGuess = val;
    else if (key == "objectGuess") r.objectGuess = val;
    else if (key == "wsServerUrl") r.wsServerUrl = val;
    else if (key == "delta_p") r.delta_p = atof(val.c_str());
    else if (key == "seed") r.seed = strtoull(val.c_str(), nullptr, 10);
    else if (key == "MLC") r.compactGroups = iv > 0;
    else {
      if (err) *err = "unknown option -" + key;
      return false;
    }
  }
  finalize();
  return true;
}

bool Parameters::setKeyword(const std::string& key, const std::string& val, std::string* err) {
  const long iv = atol(val.c_str());
  const double dv = atof(val.c_str());
  // guards as in Parameters::parseFromCPython (Parameters.cpp:380-653)
  if (key == "jobID") { if (!val.empty()) r.reconstructionID = val; }
  else if (key == "fp") { if (!val.empty()) p.dataFilePattern = val; }
  else if (key == "fs") { if (iv >= 0) p.fileStartIndex = (unsigned)iv; }
  else if (key == "hdf5path") { if (!val.empty()) p.hdf5DatasetName = val; }
  else if (key == "dpf") { if (iv > 1) p.diffsPerFile = (unsigned)iv; }
  else if (key == "beamSize") { if (dv > 0) e.beamSize = dv; }
  else if (key == "probeGuess") { if (!val.empty()) r.probeGuess = val; }
  else if (key == "objectGuess") { if (!val.empty()) r.objectGuess = val; }
  else if (key == "size") {
    if (iv > 16) r.desiredShape = (unsigned)iv;
    else { if (err) *err = "the size must be large than 16"; return false; }
  } else if (key == "qx") p.symmetric_array_center_x = (int)iv;
  else if (key == "qy") p.symmetric_array_center_y = (int)iv;
  else if (key == "nx") p.raw_file_size_x = (int)iv;
  else if (key == "ny") p.raw_file_size_y = (int)iv;
  else if (key == "scanDimsx") { if (iv > 0) e.scanDims_x = (unsigned)iv; }
  else if (key == "scanDimsy") { if (iv > 0) e.scanDims_y = (unsigned)iv; }
  else if (key == "spiralScan") { if (iv == 1) e.spiralScan = true; }
  else if (key == "flipScanAxis") { if (iv == 1) r.flipScanAxis = true; }
  else if (key == "mirror1stScanAxis") { if (iv == 1) r.mirrorX = true; }
  else if (key == "mirror2ndScanAxis") { if (iv == 1) r.mirrorY = true; }
  else if (key == "stepx") { if (dv > 0) e.stepSize_x = (float)dv; }
  else if (key == "stepy") { if (dv > 0) e.stepSize_y = (float)dv; }
  else if (key == "probeModes") { if (iv >= 1) r.probeModes = r.nProbes = (unsigned)iv; }
  else if (key == "lambd") { if (dv > 0) e.lambda = dv; }
  else if (key == "dx_d") { if (dv > 0) e.dx_d = dv; }
  else if (key == "z") { if (dv > 0) e.z_d = dv; }
  else if (key == "iter") { if (iv >= 0) r.iterations = (unsigned)iv; }
  else if (key == "T") { if (iv > 0) r.time = (double)iv; }
  else if (key == "jitterRadius") { if (iv > 0) r.jitterRadius = (unsigned)iv; }
  else if (key == "delta_p") { r.delta_p = dv; }  // the reference accepts and drops it; honoured here
  else if (key == "threshold") { if (iv > 0) p.threshold_raw_data = (float)iv; }
  else if (key == "rotate90") { if (iv >= 0) p.rotate_90_times = (unsigned)iv; }
  else if (key == "sqrtData") { if (iv == 1) p.flags |= 1; }
  else if (key == "fftShiftData") { if (iv == 1) p.flags |= 4; }
  else if (key == "binaryOutput") { if (iv == 1) r.binaryOutput = true; }
  else if (key == "simulate") { if (iv == 1) r.simulated = true; }
  else if (key == "phaseConstraint") { if (iv >= 0) r.phaseConstraint = (unsigned)iv; }
  else if (key == "updateProbe") { if (iv >= 0) r.updateProbe = (unsigned)iv; }
  else if (key == "updateModes") { if (iv >= 0) r.updateProbeModes = (unsigned)iv; }
  else if (key == "beamstopMask") { if (iv == 1) p.beamstopMask = true; }
  else if (key == "lf") { if (!val.empty()) r.positionsFilename = val; }
  else if (key == "PPS") { if (iv >= 1) r.probe_pos_search = (int)iv; }
  else if (key == "algorithm") { if (!val.empty()) r.algorithm = val; }
  else if (key == "RMS") { r.calculateRMS = iv > 0; }
  else if (key == "seed") { r.seed = strtoull(val.c_str(), nullptr, 10); }
  else if (key == "MLC") { r.compactGroups = iv > 0; }
  else {
    if (err) *err = "unknown keyword " + key;
    return false;
  }
  finalize();
  return true;
}

// ---------------------------------------------------------------- meshes
void ScanMesh::finish(const std::vector<F2>& raw, unsigned rows, unsigned cols, unsigned pitch, unsigned probe,
                      bool flip) {
  // extents over the pitched buffer with FLT_MIN / FLT_MAX identities (utilitiesKernels.cu:1291-1301);
  // raw holds rows x pitch entries, padding included
  F2 mx{FLT_MIN, FLT_MIN}, mn{FLT_MAX, FLT_MAX};
  for (const F2& v : raw) {
    mx.x = std::max(mx.x, v.x);
    mx.y = std::max(mx.y, v.y);
    mn.x = std::min(mn.x, v.x);
    mn.y = std::min(mn.y, v.y);
  }
  maxima = mx;
  minima = mn;
  canvas_x = (int)std::ceil((mx.x - mn.x) + (float)probe);
  canvas_y = (int)std::ceil((mx.y - mn.y) + (float)probe);
  positions.clear();
  indices.clear();
  for (unsigned i = 0; i < rows; ++i)
    for (unsigned j = 0; j < cols; ++j) {
      const F2 v = raw[(size_t)i * pitch + j];
      if (ml)
        positions.push_back(v);
      else
        positions.push_back(flip ? F2{v.y - mn.y, v.x - mn.x} : F2{v.x - mn.x, v.y - mn.y});
      indices.push_back(i * cols + j);
    }
  if (ml) {
    const float ext[4] = {mn.x, mn.y, mx.x, mx.y};
    int32_t canvas[2];
    ptycho_mls_canvas(&positions[0].x, (int64_t)positions.size(), ext, (int32_t)probe, canvas, box);
    canvas_x = canvas[0];
    canvas_y = canvas[1];
  }
}

bool CartesianScanMesh::generate(const Parameters& prm, std::string* err) {
  ptycho_mesh_params mp{};
  mp.nx = (int32_t)prm.e.scanDims_x;
  mp.ny = (int32_t)prm.e.scanDims_y;
  mp.step_x = prm.e.stepSize_x;
  mp.step_y = prm.e.stepSize_y;
  mp.dx_s = prm.e.dx_s;
  mp.drift_x = prm.e.drift_x;
  mp.drift_y = prm.e.drift_y;
  mp.jitter_radius = (int32_t)prm.r.jitterRadius;
  mp.mirror_x = prm.r.mirrorX;
  mp.mirror_y = prm.r.mirrorY;
  mp.flip = prm.r.flipScanAxis;
  mp.probe_size = (int32_t)prm.r.desiredShape;
  mp.seed = prm.r.seed;
  const size_t n = (size_t)mp.nx * mp.ny;
  if (ml) {
    // generateMeshML: the same mesh kernel (with dx = dx_s), extents over the pitched buffer
    // (rows padded with zeros to a multiple of 32), positions kept unshifted
    std::vector<float> raw(2 * n);
    if (ptycho_mesh_cartesian_raw(&mp, raw.data()) != 0) {
      if (err) *err = "bad Cartesian mesh parameters";
      return false;
    }
    const unsigned pitch = ((unsigned)mp.ny + 31u) / 32u * 32u;
    std::vector<F2> pitched((size_t)mp.nx * pitch, F2{0.f, 0.f});
    for (int i = 0; i < mp.nx; ++i)
      for (int j = 0; j < mp.ny; ++j)
        pitched[(size_t)i * pitch + j] = F2{raw[2 * ((size_t)i * mp.ny + j)], raw[2 * ((size_t)i * mp.ny + j) + 1]};
    finish(pitched, (unsigned)mp.nx, (unsigned)mp.ny, pitch, prm.r.desiredShape, prm.r.flipScanAxis);
    return true;
  }
  std::vector<float> pos(2 * n);
  int32_t canvas[2];
  float ext[4];
  if (ptycho_mesh_cartesian(&mp, pos.data(), canvas, ext) != 0) {
    if (err) *err = "bad Cartesian mesh parameters";
    return false;
  }
  positions.resize(n);
  indices.resize(n);
  for (size_t q = 0; q < n; ++q) {
    positions[q] = F2{pos[2 * q], pos[2 * q + 1]};
    indices[q] = (unsigned)q;
  }
  minima = F2{ext[0], ext[1]};
  maxima = F2{ext[2], ext[3]};
  canvas_x = canvas[0];
  canvas_y = canvas[1];
  return true;
}

// list and spiral meshes live in a (ceil(n/1024) x 1024) pitched buffer in the reference, and
// every entry of it -- padding included -- goes through the position kernel and the extents
static const unsigned kListPitch = 1024;

bool ListScanMesh::generate(const Parameters& prm, std::string* err) {
  const size_t n = metres.size() / 2;
  if (n == 0) {
    if (err) *err = "empty position list";
    return false;
  }
  const unsigned rows = (unsigned)((n + kListPitch - 1) / kListPitch);
  const float o = (float)(1.0 / prm.e.dx_s);
  std::mt19937_64 gen(prm.r.seed);
  std::vector<F2> raw((size_t)rows * kListPitch);
  for (unsigned i = 0; i < rows; ++i)
    for (unsigned j = 0; j < kListPitch; ++j) {
      const size_t q = (size_t)i * kListPitch + j;
      float jx = 0.f, jy = 0.f;
      if (prm.r.jitterRadius > 0) {
        const float r = (float)prm.r.jitterRadius;
        jx = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
        jy = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
      }
      float vx = q < n ? (float)metres[2 * q] : 0.f, vy = q < n ? (float)metres[2 * q + 1] : 0.f;
      vx = std::fmaf((float)i, prm.e.drift_x, vx);
      vy = std::fmaf((float)j, prm.e.drift_y, vy);
      raw[q] = F2{std::fmaf(vx, o, jx), std::fmaf(vy, o, jy)};
    }
  std::vector<F2> first(raw.begin(), raw.end());
  finish(first, 1, (unsigned)n, (unsigned)raw.size(), prm.r.desiredShape, prm.r.flipScanAxis);
  return true;
}

bool SpiralScanMesh::generate(const Parameters& prm, std::string* err) {
  const size_t n = (size_t)prm.e.scanDims_x * prm.e.scanDims_y;
  if (n == 0) {
    if (err) *err = "empty scan";
    return false;
  }
  const unsigned rows = (unsigned)((n + kListPitch - 1) / kListPitch);
  const float o = (float)(1.0 / prm.e.dx_s);
  std::mt19937_64 gen(prm.r.seed);
  std::vector<F2> raw((size_t)rows * kListPitch);
  for (unsigned i = 0; i < rows; ++i)
    for (unsigned j = 0; j < kListPitch; ++j) {
      const size_t q = (size_t)i * kListPitch + j;
      float jx = 0.f, jy = 0.f;
      if (prm.r.jitterRadius > 0) {
        const float r = (float)prm.r.jitterRadius;
        jx = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
        jy = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
      }
      const float s = std::sqrt((float)q);
      float vx = s * std::cos(4.0f * s) * prm.e.stepSize_x;
      float vy = s * std::sin(4.0f * s) * prm.e.stepSize_y;
      vx = std::fmaf((float)i, prm.e.drift_x, vx);
      vy = std::fmaf((float)j, prm.e.drift_y, vy);
      raw[q] = F2{std::fmaf(vx, o, jx), std::fmaf(vy, o, jy)};
    }
  finish(raw, 1, (unsigned)n, (unsigned)raw.size(), prm.r.desiredShape, prm.r.flipScanAxis);
  return true;
}

// ---------------------------------------------------------------- methods
real_t EPIE::iteration(Diffractions* diffs, Probe*, Sample*, IPtychoScanMesh*, std::vector<std::vector<real_t> >&,
                       const std::vector<float2>&, bool phaseConstraint, bool updateProbe, bool updateProbeModes,
                       unsigned int, bool RMS) {
  // random visiting order: repeatedly take a random one of the remaining positions
  // (ePIE.cpp:80-97, rand() % remaining; seeded here)
  const int n = diffs->n;
  std::vector<int32_t> remaining(n);
  for (int i = 0; i < n; ++i) remaining[i] = i;
  order_.resize(n);
  for (int s = 0; s < n; ++s) {
    const size_t k = (size_t)(rng_() % remaining.size());
    order_[s] = remaining[k];
    remaining.erase(remaining.begin() + k);
  }
  uint32_t flags = 0;
  if (phaseConstraint) flags |= PTYK_FLAG_PHASE_CONSTRAINT;
  if (updateProbe) flags |= PTYK_FLAG_UPDATE_PROBE;
  if (updateProbeModes) flags |= PTYK_FLAG_UPDATE_MODES;
  if (RMS) flags |= PTYK_FLAG_RMS;
  double err = 0.0;
  status_ = ptyk_epie_iteration(diffs->ctx, order_.data(), flags, &err);
  return (real_t)err;
}

real_t DMMethod::iteration(Diffractions* diffs, Probe*, Sample*, IPtychoScanMesh*, std::vector<std::vector<real_t> >&,
                           const std::vector<float2>&, bool phaseConstraint, bool updateProbe, bool, unsigned int,
                           bool) {
  uint32_t flags = 0;
  if (phaseConstraint) flags |= PTYK_FLAG_PHASE_CONSTRAINT;
  if (updateProbe) flags |= PTYK_FLAG_UPDATE_PROBE;
  status_ = ptyk_dm_iteration(diffs->ctx, flags);
  return 0;  // DM::iteration returns 0 (DM.cu:127)
}

void MLSMethod::initMem(IPtychoScanMesh* scanMesh, uint2) {
  n_ = (int64_t)scanMesh->positions.size();
  if (compact_) {
    // get_close_indices: one compact partition, its group order shuffled every iteration
    std::vector<float> pos(2 * (size_t)n_);
    for (int64_t q = 0; q < n_; ++q) {
      pos[2 * q] = scanMesh->positions[q].x;
      pos[2 * q + 1] = scanMesh->positions[q].y;
    }
    ptycho_mls_compact_groups(pos.data(), n_, seed_ + 101, nullptr, nullptr, &groups_);
    frames_.assign((size_t)n_, 0);
    offsets_.assign((size_t)groups_ + 1, 0);
    ptycho_mls_compact_groups(pos.data(), n_, seed_ + 101, frames_.data(), offsets_.data(), &groups_);
    parts_ = 1;
    return;
  }
  ptycho_mls_partitions(n_, niter_, seed_ + 101, nullptr, nullptr, &groups_, &parts_);
  frames_.assign((size_t)parts_ * n_, 0);
  offsets_.assign((size_t)parts_ * (groups_ + 1), 0);
  ptycho_mls_partitions(n_, niter_, seed_ + 101, frames_.data(), offsets_.data(), &groups_, &parts_);
}

real_t MLSMethod::iteration(Diffractions* diffs, Probe*, Sample*, IPtychoScanMesh*,
                            std::vector<std::vector<real_t> >& fourierErrors, const std::vector<float2>&, bool, bool,
                            bool, unsigned int iter, bool) {
  if (parts_ == 0) {
    status_ = PTYK_ESTATE;
    return 1;
  }
  // randNoIndex = rand_gen_no(1, miniter) - 1 and random_shuffle of the group order (MLS.cpp:106-113)
  const int32_t part = (int32_t)(rng_() % (uint64_t)parts_);
  std::vector<int32_t> order(groups_);
  for (int32_t g = 0; g < groups_; ++g) order[g] = g;
  for (int32_t i = groups_ - 1; i > 0; --i) std::swap(order[i], order[rng_() % (uint64_t)(i + 1)]);
  const int32_t* fr = frames_.data() + (size_t)part * n_;
  const int32_t* of = offsets_.data() + (size_t)part * (groups_ + 1);
  lastGroups.clear();
  lastOffsets.assign(1, 0);
  for (int32_t g : order) {
    lastGroups.insert(lastGroups.end(), fr + of[g], fr + of[g + 1]);
    lastOffsets.push_back((int32_t)lastGroups.size());
  }
  // fourierErrors[iterC].resize(diffSize) (MLS.cpp:84); the caller sizes the outer vector
  std::vector<float> fe((size_t)n_, 0.f);
  status_ = ptyk_mls_iteration(diffs->ctx, (int32_t)iter, lastGroups.data(), lastOffsets.data(), groups_, fe.data());
  if (iter >= 1) {
    if (fourierErrors.size() < iter) fourierErrors.resize(iter);
    fourierErrors[iter - 1].assign(fe.begin(), fe.end());
  }
  return 1;  // MLS::iteration returns 1 (MLS.cpp:489)
}

namespace {
// Raw frame files (FileManager.cpp:130-214): ".bin" holds consecutive frames of raw_rows x
// raw_cols float32 (BinFile: frame i at offset i x rows x cols), ".csv" comma-separated values
// read in order (CSVFile); ".h5" needs HDF5, which this build does not link.
struct FrameFile {
  std::string path;
  bool csv = false;
  std::vector<float> values;  // csv: whole file
  FILE* f = nullptr;
  ~FrameFile() {
    if (f) fclose(f);
  }
};

bool open_frame_file(const std::string& path, FrameFile* ff, std::string* err) {
  const size_t dot = path.find_last_of('.');
  if (dot == std::string::npos || dot == 0) {
    *err = "Error in -fp formatting parameter! No file extension found (" + path + ")";
    return false;
  }
  std::string ext = path.substr(dot + 1);
  std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
  ff->path = path;
  if (ext == "bin") {
    ff->f = fopen(path.c_str(), "rb");
    if (!ff->f) {
      *err = "ERROR: Failed to open file (" + path + ")";
      return false;
    }
    return true;
  }
  if (ext == "csv") {
    std::ifstream in(path);
    if (!in.is_open()) {
      *err = "ERROR: Failed to open file (" + path + ")";
      return false;
    }
    ff->csv = true;
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream iss(line);
      std::string tok;
      while (std::getline(iss, tok, ',')) {
        const char* b = tok.c_str();
        char* e = nullptr;
        const float v = std::strtof(b, &e);
        if (e != b) ff->values.push_back(v);
      }
    }
    return true;
  }
  if (ext == "h5") {
    *err = "HDF5 frame files (" + path + ") need an HDF5-enabled build; convert them to .bin or .csv";
    return false;
  }
  *err = "unsupported frame file extension (" + path + "): use .bin or .csv";
  return false;
}

bool read_frame(FrameFile* ff, unsigned section, int rows, int cols, float* out, std::string* err) {
  const size_t cnt = (size_t)rows * cols;
  if (ff->csv) {
    if ((section + 1) * cnt > ff->values.size()) {
      *err = "frame " + std::to_string(section) + " is beyond the end of " + ff->path;
      return false;
    }
    std::copy(ff->values.begin() + section * cnt, ff->values.begin() + (section + 1) * cnt, out);
    return true;
  }
  if (fseeko(ff->f, (off_t)(section * cnt * sizeof(float)), SEEK_SET) != 0 ||
      fread(out, sizeof(float), cnt, ff->f) != cnt) {
    *err = "frame " + std::to_string(section) + " is beyond the end of " + ff->path;
    return false;
  }
  return true;
}
}  // namespace

// Diffractions::load (Diffractions.cpp:188-232) and DiffractionLoader::run: frame q (scan order) is
// section q % diffsPerFile of the file the pattern names for q / diffsPerFile + fileStartIndex
// ('#' in the pattern starts a printf conversion, IPhaser.cpp:209); each raw frame is cropped,
// rotated and shifted on the GPU (ptyk_preprocess_frames) in batches.
bool IPhaser::read_frames(int n, int m, std::vector<float>* amp) {
  std::string pat = prm.p.dataFilePattern;
  const size_t hash = pat.find('#');
  if (hash != std::string::npos) pat[hash] = '%';
  const int rows = prm.p.raw_file_size_x, cols = prm.p.raw_file_size_y;
  if (rows <= 0 || cols <= 0) {
    err = "raw frame size (-nx, -ny) must be positive";
    return false;
  }
  const unsigned dpf = std::max(1u, prm.p.diffsPerFile);
  ptyk_ingest_params ip{};
  ip.raw_rows = rows;
  ip.raw_cols = cols;
  ip.center_row = prm.p.symmetric_array_center_x;
  ip.center_col = prm.p.symmetric_array_center_y;
  ip.rotate90 = (int32_t)prm.p.rotate_90_times;
  ip.fft_shift = (prm.p.flags & 4) ? 1 : 0;
  ip.threshold_sqrt = 0;  // disabled in the reference's load path (Diffractions.cpp:226-229)
  ip.square_root = (prm.p.flags & 1) ? 1 : 0;
  ip.threshold = prm.p.threshold_raw_data;
  const size_t mm = (size_t)m * m, raw_sz = (size_t)rows * cols;
  amp->assign((size_t)n * mm, 0.f);
  const int batch = 128;
  std::vector<float> raw((size_t)batch * raw_sz);
  std::unique_ptr<FrameFile> ff;
  for (int q0 = 0; q0 < n; q0 += batch) {
    const int cnt = std::min(batch, n - q0);
    for (int b = 0; b < cnt; ++b) {
      const unsigned idx = mesh->indices.empty() ? (unsigned)(q0 + b) : mesh->indices[q0 + b];
      char path[4096];
      snprintf(path, sizeof(path), pat.c_str(), (int)(idx / dpf + prm.p.fileStartIndex));
      if (!ff || ff->path != path) {
        ff.reset(new FrameFile());
        if (!open_frame_file(path, ff.get(), &err)) return false;
      }
      if (!read_frame(ff.get(), idx % dpf, rows, cols, raw.data() + (size_t)b * raw_sz, &err)) return false;
    }
    if (ptyk_preprocess_frames(eng.ctx, raw.data(), cnt, &ip, amp->data() + (size_t)q0 * mm) != PTYK_OK) {
      err = ptyk_last_error(eng.ctx);
      return false;
    }
  }
  return true;
}

bool IPhaser::simulate_frames(std::vector<float>* amp) {
  // Diffractions::simulate (Diffractions.cpp:80-116) over a seeded random-phase object of the
  // canvas size with the simulated probe (the reference reads lena/baboon PGM files here)
  const int n = eng.n, m = eng.m;
  std::mt19937_64 g(prm.r.seed + 17);
  std::vector<float> obj((size_t)2 * eng.rows * eng.cols);
  for (size_t q = 0; q < obj.size() / 2; ++q) {
    const double a = 0.9 + 0.1 * ((double)(g() >> 11) / 9007199254740992.0);
    const double ph = 0.5 * (2.0 * ((double)(g() >> 11) / 9007199254740992.0) - 1.0);
    obj[2 * q] = (float)(a * std::cos(ph));
    obj[2 * q + 1] = (float)(a * std::sin(ph));
  }
  std::vector<float> probe((size_t)2 * m * m);
  const double w = (double)m / (2 * 2.35);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const double a = i - m * 0.5, b = j - m * 0.5, r = std::sqrt(a * a + b * b);
      double v = r * prm.e.dx_s / prm.e.beamSize;
      v = v != 0 ? std::sin(M_PI * v) / (M_PI * v) : 1.0;
      probe[2 * (i * m + j)] = (float)std::fabs(v * std::exp(-r * r / (w * w)));
    }
  std::vector<float> pos(2 * (size_t)n);
  // MLs positions are centred: simulate at the MLs ROI frame, round(pos + ceil(Np/2 - M/2)) - 1
  // before rounding (the reference's simulateMLs shifts by the minima instead, a constant offset)
  const float cx = mesh->ml ? (float)std::ceil(eng.rows * 1.0 / 2 - m * 1.0 / 2) - 1.f : 0.f;
  const float cy = mesh->ml ? (float)std::ceil(eng.cols * 1.0 / 2 - m * 1.0 / 2) - 1.f : 0.f;
  for (int q = 0; q < n; ++q) {
    pos[2 * q] = mesh->positions[q].x + cx;
    pos[2 * q + 1] = mesh->positions[q].y + cy;
  }
  amp->resize((size_t)n * m * m);
  // MLs consumes intensities (its simulator's output is square-rooted again in prePhase,
  // Diffractions.cpp:150 with IPhaser.cpp:336-337; here the intensities are simulated directly)
  return ptyk_simulate(eng.ctx, obj.data(), eng.rows, eng.cols, probe.data(), pos.data(), n, mesh->ml ? 1 : 0,
                       amp->data()) == 0;
}

bool IPhaser::init() {
  prm.finalize();
  const bool mls = prm.r.algorithm == "MLs";
  // initScanMesh (IPhaser.cpp:128-171)
  if (!positions_m.empty()) {
    auto* lm = new ListScanMesh();
    lm->metres = positions_m;
    mesh.reset(lm);
  } else if (!prm.r.positionsFilename.empty()) {
    std::ifstream in(prm.r.positionsFilename);
    if (!in.is_open()) {
      err = "Failed to load raster positions file (" + prm.r.positionsFilename + ")";
      return false;
    }
    auto* lm = new ListScanMesh();
    const size_t n = (size_t)prm.e.scanDims_x * prm.e.scanDims_y;
    for (size_t i = 0; i < n; ++i) {
      float x, y;  // parsed straight into float, as ListScanMesh::loadList does (ListScanMesh.cpp:64-89)
      char comma;
      if (!(in >> x >> comma >> y) || comma != ',') {
        err = "Invalid raster positions file format [" + prm.r.positionsFilename + "]";
        return false;
      }
      lm->metres.push_back(x);
      lm->metres.push_back(y);
    }
    mesh.reset(lm);
  } else if (prm.e.spiralScan) {
    mesh.reset(new SpiralScanMesh());
  } else {
    mesh.reset(new CartesianScanMesh());
  }
  mesh->ml = mls;  // generateMeshML + get_nonoverlapping_indices + precalculate_ROI for MLs
  if (!mesh->generate(prm, &err)) return false;
  const int n = (int)mesh->positions.size(), m = (int)prm.r.desiredShape;
  if (!prm.r.simulated && diffraction.empty() && prm.p.dataFilePattern.empty()) {
    err = "Could not find data path parameter! Did you forget to provide `-fp` argument?";
    return false;
  }
  if (!diffraction.empty() && (diff_n != n || diff_rows != m || diff_cols != m)) {
    err = "diffraction array must be N x size x size with N = number of scan positions";
    return false;
  }
  // initSample / initProbe / initDiffractions: one engine handle
  ptyk_config cfg{};
  cfg.n_frames = n;
  cfg.frame_size = m;
  cfg.n_modes = (int32_t)prm.r.probeModes;
  cfg.obj_rows = mesh->canvas_x;
  cfg.obj_cols = mesh->canvas_y;
  cfg.device = 0;
  if (eng.ctx) {
    ptyk_destroy(eng.ctx);
    eng.ctx = nullptr;
  }
  if (ptyk_create(&cfg, &eng.ctx) != PTYK_OK) {
    err = ptyk_last_error(nullptr);
    return false;
  }
  eng.n = n;
  eng.m = m;
  eng.k = cfg.n_modes;
  eng.rows = cfg.obj_rows;
  eng.cols = cfg.obj_cols;
  eng.bind();
  if (!mls) {  // MLs positions stay centred; their ROIs are set up by ptyk_mls_begin
    std::vector<float> pos(2 * (size_t)n);
    for (int q = 0; q < n; ++q) {
      pos[2 * q] = mesh->positions[q].x;
      pos[2 * q + 1] = mesh->positions[q].y;
    }
    if (ptyk_set_positions(eng.ctx, pos.data()) != PTYK_OK) {
      err = ptyk_last_error(eng.ctx);
      return false;
    }
  }
  std::vector<float> amp;
  if (!diffraction.empty()) {
    amp = diffraction;
  } else if (!prm.r.simulated) {
    if (!read_frames(n, m, &amp)) return false;
  } else if (!simulate_frames(&amp)) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (mls) {
    // MLs takes intensities: per-frame sums (fillSquaredSums) then the square root
    // (squaredRoot), IPhaser.cpp:336-337
    const size_t mm = (size_t)m * m;
    frame_power_.assign(n, 0.0);
    for (int q = 0; q < n; ++q) {
      double sp = 0.0;
      for (size_t i = 0; i < mm; ++i) {
        float& v = amp[(size_t)q * mm + i];
        if (!(v >= 0.f)) {
          err = "MLs diffraction intensities must be >= 0";
          return false;
        }
        sp += v;
        v = std::sqrt(v);
      }
      if (!(sp > 0.0)) {
        err = "MLs diffraction frame " + std::to_string(q) + " is all zero";
        return false;
      }
      frame_power_[q] = sp;
    }
  }
  if (ptyk_set_frames_f32(eng.ctx, amp.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  return true;
}

bool IPhaser::addPhasingMethod(const std::string& name, unsigned iterations) {
  if (name == "ePIE")
    methods.push_back(CXPhasing{name, std::unique_ptr<IPhasingMethod>(new EPIE(prm.r.seed)), iterations});
  else if (name == "DM")
    methods.push_back(CXPhasing{name, std::unique_ptr<IPhasingMethod>(new DMMethod()), iterations});
  else if (name == "MLs")
    methods.push_back(
        CXPhasing{name, std::unique_ptr<IPhasingMethod>(new MLSMethod(prm.r.seed, prm.r.Niter, prm.r.compactGroups)), iterations});
  else {
    err = "Reconstruction algorithm [" + name + "] not implemented";
    return false;
  }
  return true;
}

namespace {
// The object guess into the canvas.  A guess read from a file (Sample::loadGuess ->
// h_setObjectArray, Sample.cpp:164-182, SampleKernels.cu:311-326 and 626-640) is centred: its
// element (i, j) lands at (i + rows/2 - r/2, j + cols/2 - c/2), halves in integer division.  A
// guess given as an array (objectNP, Sample::loadarr) fills the canvas from its origin.  Elements
// that would land outside the canvas are dropped (the reference's unsigned offsets write outside
// it for a guess larger than the canvas).
void place_object_guess(std::vector<float>* canvas, int rows, int cols, const std::vector<float>& g, int r, int c,
                        bool centred) {
  const int oi = centred ? rows / 2 - r / 2 : 0, oj = centred ? cols / 2 - c / 2 : 0;
  for (int i = 0; i < r; ++i) {
    const int ti = i + oi;
    if (ti < 0 || ti >= rows) continue;
    for (int j = 0; j < c; ++j) {
      const int tj = j + oj;
      if (tj < 0 || tj >= cols) continue;
      (*canvas)[2 * ((size_t)ti * cols + tj)] = g[2 * ((size_t)i * c + j)];
      (*canvas)[2 * ((size_t)ti * cols + tj) + 1] = g[2 * ((size_t)i * c + j) + 1];
    }
  }
}
}  // namespace

bool IPhaser::prePhase() {
  if (!eng.ctx) {
    err = "init() must be called before phase()";
    return false;
  }
  if (mesh->ml) return mls_prephase();
  // IPhaser::prePhase, non-MLs branch (IPhaser.cpp:296-333)
  const int m = eng.m;
  if (!prm.r.probeGuess.empty()) {
    std::vector<float> g;
    int r = 0, c = 0;
    if (!load_complex_csv(prm.r.probeGuess, &g, &r, &c) || r != m || c != m) {
      err = "Probe guess file (" + prm.r.probeGuess + ") not found or not " + std::to_string(m) + "x" +
            std::to_string(m);
      return false;
    }
    probe_guess = g;
    probe_guess_rows = r;
    probe_guess_cols = c;
  }
  if (!probe_guess.empty()) {
    if (probe_guess_rows != m || probe_guess_cols != m) {
      err = "probe guess must be size x size";
      return false;
    }
    std::vector<float> all((size_t)2 * eng.k * m * m, 0.f);
    std::copy(probe_guess.begin(), probe_guess.end(), all.begin());
    if (ptyk_set_probe(eng.ctx, all.data()) != PTYK_OK || ptyk_refresh_maxima(eng.ctx) != PTYK_OK) {
      err = ptyk_last_error(eng.ctx);
      return false;
    }
  } else if (ptyk_init_probe(eng.ctx, prm.e.beamSize, prm.e.dx_s) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (ptyk_clear_object(eng.ctx) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (!prm.r.objectGuess.empty()) {
    std::vector<float> g;
    int r = 0, c = 0;
    if (!load_complex_csv(prm.r.objectGuess, &g, &r, &c)) {
      err = "Object guess file (" + prm.r.objectGuess + ") not found";
      return false;
    }
    object_guess = g;
    obj_guess_rows = r;
    obj_guess_cols = c;
    obj_guess_centred = true;
  }
  if (!object_guess.empty()) {
    std::vector<float> canvas((size_t)2 * eng.rows * eng.cols, 0.f);
    place_object_guess(&canvas, eng.rows, eng.cols, object_guess, obj_guess_rows, obj_guess_cols, obj_guess_centred);
    if (ptyk_set_object(eng.ctx, canvas.data()) != PTYK_OK) {
      err = ptyk_last_error(eng.ctx);
      return false;
    }
  }
  errors.assign(prm.r.iterations, 0.0);
  fourierErrors.assign(prm.r.iterations, std::vector<real_t>());  // IPhaser.cpp:341
  for (auto& pm : methods) pm.method->initMem(mesh.get(), uint2{(unsigned)m, (unsigned)m});
  return true;
}

namespace {
// Standard normal draws (Box-Muller over the seeded uniform stream); the reference's
// normalRandom is time-seeded.
struct Normal {
  std::mt19937_64 g;
  bool have = false;
  double spare = 0.0;
  explicit Normal(uint64_t seed) : g(seed) {}
  double operator()() {
    if (have) {
      have = false;
      return spare;
    }
    const double u1 = (double)uniform_open0(g), u2 = (double)uniform_open0(g);
    const double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(2.0 * M_PI * u2);
    have = true;
    return r * std::cos(2.0 * M_PI * u2);
  }
};

// Eigen-decomposition of a Hermitian k x k matrix by cyclic complex Jacobi rotations (FP64):
// a (row-major, destroyed) -> eigenvalues w, eigenvectors as the columns of v.
void hermitian_eig(std::vector<std::complex<double>>& a, int k, std::vector<double>* w,
                   std::vector<std::complex<double>>* v) {
  typedef std::complex<double> cd;
  v->assign((size_t)k * k, cd(0));
  for (int i = 0; i < k; ++i) (*v)[(size_t)i * k + i] = 1.0;
  auto A = [&](int r, int c) -> cd& { return a[(size_t)r * k + c]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < k; ++p)
      for (int q = 0; q < k; ++q) (p == q ? dia : off) += std::norm(A(p, q));
    if (off <= 1e-30 * (dia > 0 ? dia : 1.0)) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const cd g = A(p, q);
        if (std::abs(g) == 0.0) continue;
        // phase column/row q so that A(p, q) becomes |g|, then a real rotation of (p, q)
        const cd e = std::conj(g) / std::abs(g);
        for (int r = 0; r < k; ++r) A(r, q) *= e;
        for (int c = 0; c < k; ++c) A(q, c) *= std::conj(e);
        for (int r = 0; r < k; ++r) (*v)[(size_t)r * k + q] *= e;
        const double aa = A(p, p).real(), bb = A(q, q).real(), rr = A(p, q).real();
        const double th = 0.5 * std::atan2(-2.0 * rr, bb - aa);
        const double c = std::cos(th), s = std::sin(th);
        for (int r = 0; r < k; ++r) {
          const cd x = A(r, p), y = A(r, q);
          A(r, p) = c * x + s * y;
          A(r, q) = -s * x + c * y;
        }
        for (int cc = 0; cc < k; ++cc) {
          const cd x = A(p, cc), y = A(q, cc);
          A(p, cc) = c * x + s * y;
          A(q, cc) = -s * x + c * y;
        }
        for (int r = 0; r < k; ++r) {
          const cd x = (*v)[(size_t)r * k + p], y = (*v)[(size_t)r * k + q];
          (*v)[(size_t)r * k + p] = c * x + s * y;
          (*v)[(size_t)r * k + q] = -s * x + c * y;
        }
      }
  }
  w->resize(k);
  for (int i = 0; i < k; ++i) (*w)[i] = A(i, i).real();
}
}  // namespace

// Probe::initMLH (Probe.cpp:376-494) given its random draws: mode 0 is the given one when it is
// non-zero, else the
// Fresnel-zone-plate model (h_initProbeMLH / d_initProbeMLH / h_endProbeMLH, ProbeKernels.cu:220-366,
// 592-641): the aperture D_FZP/2 >= r >= D_H/2 times the lens chirp exp(i fzp pi r^2) and the
// propagator exp(i k r^2 / 2z), placed fftshifted, forward FFT (on the engine), times the far-field
// factor exp(i k z) exp(i k u^2 / 2z), shifted back.  The phase arguments are evaluated in double:
// the reference forms them in float (up to ~3e-4 rad of rounding at the zone-plate edge, in an order
// its compiler's contractions decide).  Modes 1..K-1 (unless all were given) are mode 0 tilted by
// exp(-2 pi i (r1 (c/M - 0.5) + r2 (r/M - 0.5))) (h_initVarProbe, ProbeKernels.cu:756-772;
// d_imshift_fft, utilitiesKernels.cu:784-826) with (r1, r2) = draws[2(i-1)], draws[2(i-1)+1]; then
// all modes are replaced by the eigenmodes of A_ij = sum conj(P_i) P_j, strongest |eigenvalue| first,
// each eigenvector's phase fixed so its largest component is real and positive.
int init_mlh_modes(ptyk_ctx* ctx, int m, int k, double lambda, double dx_s, double beam_size, const float* draws,
                   bool tilt, std::vector<float>* modes, std::string* err) {
  typedef std::complex<double> cd;
  const size_t mm = (size_t)m * m;
  if (modes->size() != 2 * mm * k) modes->assign(2 * mm * k, 0.f);
  if (!draws && tilt && k > 1) {
    *err = "draws required";
    return PTYK_EINVAL;
  }
  bool have_mode0 = false;
  for (size_t q = 0; q < 2 * mm && !have_mode0; ++q) have_mode0 = (*modes)[q] != 0.f;
  if (!have_mode0) {
    const double Rn = 90e-6, dRn = 50e-9, fl = 2 * Rn * dRn / lambda;
    const double D_FZP = 180e-6, D_H = 60e-6;
    const double dx_fzp = lambda * fl / (m * dx_s);
    const double fzp = (-2 / (lambda * 2 * fl)) * M_PI;
    const double step = dx_fzp * m / (m - 1), start = -dx_fzp * m / 2;
    const double kk = 2 * M_PI / lambda, z = fl + beam_size;
    const cd kz = std::polar(1.0, kk * z);
    const double ekdz = kk / (2 * z);
    const double fu = lambda * (fl + beam_size) / dx_fzp;
    const double fustep = fu / (m - 1), startfu = -fu / 2 + fu / (m - 1) * (m / 2);
    const double rf = D_FZP / 2, rh = D_H / 2;
    std::vector<cd> pf(mm);
    std::vector<float> f(2 * mm);
    for (int b = 0; b < m; ++b)
      for (int t = 0; t < m; ++t) {
        const double x = start + step * t, y = start + step * b;
        const double r2 = x * x + y * y, rr = std::sqrt(r2);
        const double ap = (rr <= rf && rr >= rh) ? 1.0 : 0.0;
        const double lx = t < m / 2 ? startfu + fustep * t : -fu / 2 + fustep * (t - m / 2);
        const double ly = b < m / 2 ? startfu + fustep * b : -fu / 2 + fustep * (b - m / 2);
        pf[(size_t)b * m + t] = kz * std::polar(1.0, ekdz * (lx * lx + ly * ly));
        const cd kv = ap * std::polar(1.0, fzp * r2 + ekdz * r2);
        const size_t d = (size_t)((b + m / 2) % m) * m + (t + m / 2) % m;  // fftshift placement
        f[2 * d] = (float)kv.real();
        f[2 * d + 1] = (float)kv.imag();
      }
    const int rc = ptyk_fft2(ctx, f.data(), 1, 0);
    if (rc != PTYK_OK) {
      *err = ptyk_last_error(ctx);
      return rc;
    }
    for (int b = 0; b < m; ++b)
      for (int t = 0; t < m; ++t) {
        const size_t q = (size_t)b * m + t;
        const cd v = cd(f[2 * q], f[2 * q + 1]) * pf[q];
        const size_t d = (size_t)((b + m / 2) % m) * m + (t + m / 2) % m;
        (*modes)[2 * d] = (float)v.real();
        (*modes)[2 * d + 1] = (float)v.imag();
      }
  }
  if (tilt)
    for (int i = 1; i < k; ++i) {
      const double r1 = draws[2 * (i - 1)], r2 = draws[2 * (i - 1) + 1];
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) {
          const double ph = -2 * M_PI * (r1 * ((double)c / m - 0.5) + r2 * ((double)r / m - 0.5));
          const size_t q = (size_t)r * m + c;
          const cd v = cd((*modes)[2 * q], (*modes)[2 * q + 1]) * std::polar(1.0, ph);
          (*modes)[2 * (i * mm + q)] = (float)v.real();
          (*modes)[2 * (i * mm + q) + 1] = (float)v.imag();
        }
    }
  if (k > 1) {
    std::vector<cd> a((size_t)k * k), v;
    std::vector<double> w;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        cd s = 0;
        for (size_t q = 0; q < mm; ++q)
          s += std::conj(cd((*modes)[2 * (i * mm + q)], (*modes)[2 * (i * mm + q) + 1])) *
               cd((*modes)[2 * (j * mm + q)], (*modes)[2 * (j * mm + q) + 1]);
        a[(size_t)i * k + j] = s;
      }
    hermitian_eig(a, k, &w, &v);
    std::vector<int> ord(k);
    for (int i = 0; i < k; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return std::fabs(w[x]) > std::fabs(w[y]); });
    std::vector<float> out(modes->size());
    for (int o = 0; o < k; ++o) {
      const int e = ord[o];
      int big = 0;
      for (int r = 1; r < k; ++r)
        if (std::abs(v[(size_t)r * k + e]) > std::abs(v[(size_t)big * k + e])) big = r;
      const cd ph = std::conj(v[(size_t)big * k + e]) / std::abs(v[(size_t)big * k + e]);
      for (size_t q = 0; q < mm; ++q) {
        cd s = 0;
        for (int r = 0; r < k; ++r)
          s += cd((*modes)[2 * (r * mm + q)], (*modes)[2 * (r * mm + q) + 1]) * (v[(size_t)r * k + e] * ph);
        out[2 * (o * mm + q)] = (float)s.real();
        out[2 * (o * mm + q) + 1] = (float)s.imag();
      }
    }
    *modes = out;
  }
  return PTYK_OK;
}

namespace {
// the MLs initial-mode draws: two standard normals per extra mode (the reference's time-seeded
// normalRandom, Probe.cpp:428-432)
std::vector<float> mlh_draws(uint64_t seed, int k) {
  Normal nrm(seed + 303);
  std::vector<float> d(2 * (size_t)std::max(k - 1, 0));
  for (auto& x : d) x = (float)nrm();
  return d;
}
}  // namespace

bool IPhaser::mls_probe_modes(std::vector<float>* modes) {
  const int m = eng.m, k = eng.k;
  const size_t mm = (size_t)m * m;
  modes->assign(2 * mm * k, 0.f);
  if (!prm.r.probeGuess.empty()) {
    std::vector<float> g;
    int r = 0, c = 0;
    if (!load_complex_csv(prm.r.probeGuess, &g, &r, &c) || r != m || c != m) {
      err = "Probe guess file (" + prm.r.probeGuess + ") not found or not " + std::to_string(m) + "x" +
            std::to_string(m);
      return false;
    }
    std::copy(g.begin(), g.end(), modes->begin());
  } else if (!probe_guess.empty()) {
    // probeNP: one mode (size x size) as in the reference, or all K modes stacked (K*size x size)
    if (probe_guess_cols != m || (probe_guess_rows != m && probe_guess_rows != k * m)) {
      err = "probe guess must be size x size or (probeModes*size) x size";
      return false;
    }
    std::copy(probe_guess.begin(), probe_guess.end(), modes->begin());
  }
  const bool all_given = !probe_guess.empty() && prm.r.probeGuess.empty() && probe_guess_rows == k * m;
  const std::vector<float> draws = mlh_draws(prm.r.seed, k);
  return init_mlh_modes(eng.ctx, m, k, prm.e.lambda, prm.e.dx_s, prm.e.beamSize, draws.data(), !all_given, modes,
                        &err) == PTYK_OK;
}

// IPhaser::prePhase, MLs branch (IPhaser.cpp:283-295, 318-358): probe modes, object (guess or
// Sample::initObject's random start), the variable mode (Probe::initVarModes) and the evolution
// draws, then Probe::initEvo in the engine.
bool IPhaser::mls_prephase() {
  const int m = eng.m, n = eng.n;
  std::vector<float> modes;
  if (!mls_probe_modes(&modes)) return false;
  if (ptyk_set_probe(eng.ctx, modes.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (!prm.r.objectGuess.empty()) {
    std::vector<float> g;
    int r = 0, c = 0;
    if (!load_complex_csv(prm.r.objectGuess, &g, &r, &c)) {
      err = "Object guess file (" + prm.r.objectGuess + ") not found";
      return false;
    }
    object_guess = g;
    obj_guess_rows = r;
    obj_guess_cols = c;
    obj_guess_centred = true;
  }
  std::vector<float> canvas((size_t)2 * eng.rows * eng.cols, 0.f);
  if (!object_guess.empty()) {
    place_object_guess(&canvas, eng.rows, eng.cols, object_guess, obj_guess_rows, obj_guess_cols, obj_guess_centred);
  } else {
    // d_initRandObjectArray (SampleKernels.cu:328-339): |O| = u1, arg O = 2 pi u2 / 10.  The
    // simulated case starts here too (the reference starts it from O = 0, which its own
    // normalisation turns into NaN, Probe.cpp:579-605)
    std::mt19937_64 g(prm.r.seed + 404);
    for (size_t q = 0; q < (size_t)eng.rows * eng.cols; ++q) {
      const float u1 = uniform_open0(g), u2 = uniform_open0(g);
      const float a = (float)(2 * M_PI * u2 * 0.1);
      canvas[2 * q] = u1 * std::cos(a);
      canvas[2 * q + 1] = u1 * std::sin(a);
    }
  }
  if (ptyk_set_object(eng.ctx, canvas.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  // d_initVarMOdes (ProbeKernels.cu:99-110): u1 exp(2 pi i u2)
  std::vector<float> vmode(2 * (size_t)m * m);
  {
    std::mt19937_64 g(prm.r.seed + 505);
    for (size_t q = 0; q < (size_t)m * m; ++q) {
      const float u1 = uniform_open0(g), u2 = uniform_open0(g);
      const float a = (float)(2 * M_PI * u2);
      vmode[2 * q] = u1 * std::cos(a);
      vmode[2 * q + 1] = u1 * std::sin(a);
    }
  }
  std::vector<double> evo(n);
  Normal nrm(prm.r.seed + 606);
  for (int i = 0; i < n; ++i) evo[i] = 1e-6 * nrm();
  std::vector<float> pos(2 * (size_t)n);
  for (int q = 0; q < n; ++q) {
    pos[2 * q] = mesh->positions[q].x;
    pos[2 * q + 1] = mesh->positions[q].y;
  }
  ptyk_mls_config cfg{};
  cfg.probe_pos_search = prm.r.probe_pos_search;
  cfg.probe_reconstruct = prm.r.probe_reconstruct;
  cfg.object_reconstruct = prm.r.object_reconstruct;
  cfg.variable_probe = prm.r.variable_probe;
  cfg.variable_intensity = prm.r.variable_intensity;
  cfg.variable_probe_modes = prm.r.variable_probe_modes;
  cfg.apply_subpix_shift = prm.r.apply_subpix_shift;
  cfg.apply_multimodal_update = prm.r.apply_multimodal_update;
  cfg.n_objects = (int32_t)prm.r.Nobjects;
  cfg.n_rec = prm.r.Nrec;
  for (int i = 0; i < 4; ++i) cfg.object_box[i] = mesh->box[i];
  cfg.delta_p = prm.r.delta_p;
  cfg.beta_lsq = prm.r.beta_LSQ;
  cfg.beta_object = prm.r.beta_object;
  cfg.beta_probe = prm.r.beta_probe;
  if (ptyk_mls_begin(eng.ctx, &cfg, pos.data(), vmode.data(), evo.data(), frame_power_.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  errors.assign(prm.r.iterations, 0.0);
  fourierErrors.assign(prm.r.iterations, std::vector<real_t>());  // IPhaser.cpp:341
  for (auto& pm : methods) pm.method->initMem(mesh.get(), uint2{(unsigned)m, (unsigned)m});
  return true;
}

bool IPhaser::phaseStep(IPhasingMethod* m, unsigned i, real_t* e) {
  // IPhaser::phaseStep (IPhaser.cpp:391-398)
  *e = m->iteration(&eng.diffractions, &eng.probe, &eng.sample, mesh.get(), fourierErrors, mesh->positions,
                    i <= prm.r.phaseConstraint, i >= prm.r.updateProbe, i >= prm.r.updateProbeModes, i,
                    prm.r.calculateRMS);
  if (m->status() != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  return true;
}

void IPhaser::postPhase() {
  if (eng.ctx) ptyk_end_modal(eng.ctx);
  for (auto& pm : methods) pm.method->endPhasing();
}

bool IPhaser::phase() {
  if (!eng.ctx || !mesh) {
    err = "init() must be called before phase()";
    return false;
  }
  t0_ = now_s();
  if (!prePhase()) return false;
  // IPhaser::phaseLoop (IPhaser.cpp:367-389)
  for (auto& pm : methods)
    for (unsigned i = 1; i <= pm.iterations; ++i) {
      real_t e = 0;
      if (!phaseStep(pm.method.get(), i, &e)) return false;
      if (i - 1 < errors.size()) errors[i - 1] += e;
      if (i % prm.r.save == 0) writeResultsToDisk((int)(i / prm.r.save));
      if (prm.r.time >= 0 && now_s() - t0_ >= prm.r.time) break;
    }
  postPhase();
  phase_seconds = now_s() - t0_;
  return true;
}

bool IPhaser::phaseinit() {
  if (!eng.ctx || !mesh) {
    err = "init() must be called before phaseinit()";
    return false;
  }
  t0_ = now_s();
  return prePhase();
}

bool IPhaser::phasestepvis(unsigned i) {
  for (auto& pm : methods) {
    real_t e = 0;
    if (!phaseStep(pm.method.get(), i, &e)) return false;
    if (i >= 1 && i - 1 < errors.size()) errors[i - 1] += e;
    if (prm.r.time >= 0 && now_s() - t0_ >= prm.r.time) break;
  }
  return true;
}

bool IPhaser::phasepost() {
  postPhase();
  phase_seconds = now_s() - t0_;
  return true;
}

bool IPhaser::writeResultsToDisk(int r) {
  if (!eng.ctx) return false;
  mkdir("data", 0755);
  const bool bin = prm.r.binaryOutput;
  std::vector<float> probe((size_t)2 * eng.k * eng.m * eng.m), obj((size_t)2 * eng.rows * eng.cols);
  if (ptyk_get_probe(eng.ctx, probe.data()) != PTYK_OK || ptyk_get_object(eng.ctx, obj.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  char fn[512];
  snprintf(fn, sizeof fn, "data/%s_probes_%d.%s", prm.r.reconstructionID.c_str(), r, bin ? "bin" : "csv");
  bool ok = save_complex(fn, probe.data(), eng.k * eng.m, eng.m, bin);
  snprintf(fn, sizeof fn, "data/%s_object_%d.%s", prm.r.reconstructionID.c_str(), r, bin ? "bin" : "csv");
  ok = save_complex(fn, obj.data(), eng.rows, eng.cols, bin) && ok;
  return ok;
}

// ---------------------------------------------------------------- CSV
bool save_complex(const std::string& path, const float* c, int rows, int cols, bool binary) {
  std::ofstream out(path, binary ? std::ios::out | std::ios::binary : std::ios::out);
  if (!out.is_open()) return false;
  if (binary) {
    out.write((const char*)c, (std::streamsize)rows * cols * 2 * sizeof(float));
    return true;
  }
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) {
      const float re = c[2 * ((size_t)i * cols + j)], im = c[2 * ((size_t)i * cols + j) + 1];
      out << re;
      if (im >= 0) out << "+";
      out << im << "j";
      if (j < cols - 1) out << ", ";
    }
    out << std::endl;
  }
  return true;
}

// Complex token "a+bj" / "a-bj" / "(a+bj)" / "a", as written by save_complex and numpy.
// The reference's operator>> (Cuda2DArray.cpp:54-102) splits at the first 'e' and assumes a
// two-digit exponent on the real part, which misreads e.g. "0.5+2.3e-06j"; here the split is
// the last sign that does not follow an exponent marker.
static void parse_complex_token(std::string t, float* re, float* im) {
  std::string u;
  for (char ch : t)
    if (ch != ',' && ch != '(' && ch != ')' && ch != ' ') u += ch;
  const bool imag = !u.empty() && (u.back() == 'j' || u.back() == 'i');
  if (imag) u.pop_back();
  size_t split = std::string::npos;
  for (size_t k = u.size(); k-- > 1;)
    if ((u[k] == '+' || u[k] == '-') && u[k - 1] != 'e' && u[k - 1] != 'E') {
      split = k;
      break;
    }
  if (split == std::string::npos) {
    if (imag) {
      *re = 0.f;
      *im = (float)atof(u.c_str());
    } else {
      *re = (float)atof(u.c_str());
      *im = 0.f;
    }
    return;
  }
  *re = (float)atof(u.substr(0, split).c_str());
  *im = (float)atof(u.substr(split).c_str());
}

bool load_complex_csv(const std::string& path, std::vector<float>* c, int* rows, int* cols) {
  std::ifstream in(path);
  if (!in.is_open()) return false;
  c->clear();
  std::string line;
  int r = 0, cc = -1;
  while (std::getline(in, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::istringstream ls(line);
    std::string tok;
    int k = 0;
    while (ls >> tok) {
      float re, im;
      parse_complex_token(tok, &re, &im);
      c->push_back(re);
      c->push_back(im);
      ++k;
    }
    if (cc < 0) cc = k;
    if (k != cc) return false;
    ++r;
  }
  *rows = r;
  *cols = cc < 0 ? 0 : cc;
  return r > 0;
}

}  // namespace ptycho
