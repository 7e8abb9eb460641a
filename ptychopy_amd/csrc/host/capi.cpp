// C ABI of the host driver (include/ptycho.h): an IPhaser behind an opaque handle.
#include <algorithm>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "phaser.h"
#include "ptycho.h"

struct ptycho_phaser {
  ptycho::IPhaser ph;
};

static thread_local std::string g_err;

extern "C" {

int ptycho_phaser_create(ptycho_phaser** out) {
  if (!out) return -1;
  *out = new ptycho_phaser();
  return 0;
}

void ptycho_phaser_destroy(ptycho_phaser* h) { delete h; }

const char* ptycho_phaser_last_error(const ptycho_phaser* h) { return h ? h->ph.err.c_str() : g_err.c_str(); }

int ptycho_phaser_set_param(ptycho_phaser* h, const char* key, const char* value) {
  if (!h || !key || !value) return -1;
  return h->ph.prm.setKeyword(key, value, &h->ph.err) ? 0 : -1;
}

int ptycho_phaser_parse_cmdline(ptycho_phaser* h, const char* cmdline) {
  if (!h || !cmdline) return -1;
  std::istringstream ss(cmdline);
  std::vector<std::string> args;
  std::string t;
  while (ss >> t) args.push_back(t);
  return h->ph.prm.parseCommandLine(args, &h->ph.err) ? 0 : -1;
}

int ptycho_phaser_set_diffraction(ptycho_phaser* h, const float* data, int64_t n, int32_t rows, int32_t cols) {
  if (!h || !data || n <= 0 || rows <= 0 || cols <= 0) return -1;
  h->ph.diffraction.assign(data, data + (size_t)n * rows * cols);
  h->ph.diff_n = (int)n;
  h->ph.diff_rows = rows;
  h->ph.diff_cols = cols;
  return 0;
}

int ptycho_phaser_set_positions(ptycho_phaser* h, const double* xy, int64_t n) {
  if (!h || !xy || n <= 0) return -1;
  h->ph.positions_m.assign(xy, xy + 2 * (size_t)n);
  return 0;
}

int ptycho_phaser_set_object_guess(ptycho_phaser* h, const float* c64, int32_t rows, int32_t cols) {
  if (!h || !c64 || rows <= 0 || cols <= 0) return -1;
  h->ph.object_guess.assign(c64, c64 + 2 * (size_t)rows * cols);
  h->ph.obj_guess_rows = rows;
  h->ph.obj_guess_cols = cols;
  h->ph.obj_guess_centred = false;
  return 0;
}

int ptycho_phaser_set_probe_guess(ptycho_phaser* h, const float* c64, int32_t rows, int32_t cols) {
  if (!h || !c64 || rows <= 0 || cols <= 0) return -1;
  h->ph.probe_guess.assign(c64, c64 + 2 * (size_t)rows * cols);
  h->ph.probe_guess_rows = rows;
  h->ph.probe_guess_cols = cols;
  return 0;
}

int ptycho_phaser_init(ptycho_phaser* h) { return h && h->ph.init() ? 0 : -1; }
int ptycho_phaser_add_method(ptycho_phaser* h, const char* name, uint32_t iterations) {
  return h && name && h->ph.addPhasingMethod(name, iterations) ? 0 : -1;
}
int ptycho_phaser_phase(ptycho_phaser* h) { return h && h->ph.phase() ? 0 : -1; }
int ptycho_phaser_phaseinit(ptycho_phaser* h) { return h && h->ph.phaseinit() ? 0 : -1; }
int ptycho_phaser_phasestep(ptycho_phaser* h, uint32_t i) { return h && h->ph.phasestepvis(i) ? 0 : -1; }
int ptycho_phaser_phasepost(ptycho_phaser* h) { return h && h->ph.phasepost() ? 0 : -1; }
int ptycho_phaser_write_results(ptycho_phaser* h, int32_t r) { return h && h->ph.writeResultsToDisk(r) ? 0 : -1; }

int ptycho_phaser_get_param(ptycho_phaser* h, const char* key, char* buf, int32_t len) {
  if (!h || !key || !buf || len <= 0) return -1;
  const auto& p = h->ph.prm;
  std::ostringstream o;
  o.precision(17);
  const std::string k = key;
  if (k == "dx_s") o << p.e.dx_s;
  else if (k == "iterations") o << p.r.iterations;
  else if (k == "size") o << p.r.desiredShape;
  else if (k == "algorithm") o << p.r.algorithm;
  else if (k == "jobID") o << p.r.reconstructionID;
  else if (k == "probeModes") o << p.r.probeModes;
  else if (k == "stepx") o << p.e.stepSize_x;
  else if (k == "stepy") o << p.e.stepSize_y;
  else if (k == "scanDimsx") o << p.e.scanDims_x;
  else if (k == "scanDimsy") o << p.e.scanDims_y;
  else if (k == "beamSize") o << p.e.beamSize;
  else if (k == "lambda") o << p.e.lambda;
  else if (k == "z") o << p.e.z_d;
  else if (k == "dx_d") o << p.e.dx_d;
  else if (k == "updateProbe") o << p.r.updateProbe;
  else if (k == "updateModes") o << p.r.updateProbeModes;
  else if (k == "phaseConstraint") o << p.r.phaseConstraint;
  else if (k == "delta_p") o << p.r.delta_p;
  else if (k == "PPS") o << p.r.probe_pos_search;
  else if (k == "RMS") o << (p.r.calculateRMS ? 1 : 0);
  else if (k == "time") o << p.r.time;
  else {
    h->ph.err = "unknown parameter " + k;
    return -1;
  }
  const std::string s = o.str();
  std::strncpy(buf, s.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return 0;
}

int ptycho_phaser_object_shape(ptycho_phaser* h, int32_t* rows, int32_t* cols) {
  if (!h || !rows || !cols || !h->ph.eng.ctx) return -1;
  *rows = h->ph.eng.rows;
  *cols = h->ph.eng.cols;
  return 0;
}

int ptycho_phaser_get_object(ptycho_phaser* h, float* out) {
  if (!h || !out || !h->ph.eng.ctx) return -1;
  return ptyk_get_object(h->ph.eng.ctx, out) == PTYK_OK ? 0 : -1;
}

int ptycho_phaser_probe_shape(ptycho_phaser* h, int32_t* modes, int32_t* size) {
  if (!h || !modes || !size || !h->ph.eng.ctx) return -1;
  *modes = h->ph.eng.k;
  *size = h->ph.eng.m;
  return 0;
}

int ptycho_phaser_get_probe(ptycho_phaser* h, float* out) {
  if (!h || !out || !h->ph.eng.ctx) return -1;
  return ptyk_get_probe(h->ph.eng.ctx, out) == PTYK_OK ? 0 : -1;
}

int ptycho_phaser_get_positions(ptycho_phaser* h, float* out, int64_t n) {
  if (!h || !out || !h->ph.mesh || (int64_t)h->ph.mesh->positions.size() != n) return -1;
  for (int64_t q = 0; q < n; ++q) {
    out[2 * q] = h->ph.mesh->positions[q].x;
    out[2 * q + 1] = h->ph.mesh->positions[q].y;
  }
  return 0;
}

int64_t ptycho_phaser_num_positions(ptycho_phaser* h) {
  return h && h->ph.mesh ? (int64_t)h->ph.mesh->positions.size() : -1;
}

int ptycho_phaser_get_errors(ptycho_phaser* h, double* out, int32_t n) {
  if (!h || !out) return -1;
  for (int32_t i = 0; i < n; ++i) out[i] = i < (int32_t)h->ph.errors.size() ? h->ph.errors[i] : 0.0;
  return 0;
}

int ptycho_phaser_get_fourier_errors(ptycho_phaser* h, uint32_t iter, float* out, int64_t n) {
  if (!h || !out || iter == 0) return -1;
  bool ml = false;
  for (auto& pm : h->ph.methods) ml = ml || dynamic_cast<ptycho::MLSMethod*>(pm.method.get()) != nullptr;
  if (!ml) {
    h->ph.err = "no MLs method";
    return -1;
  }
  const auto& fe = h->ph.fourierErrors;
  if (iter > fe.size() || (int64_t)fe[iter - 1].size() != n) return -1;
  std::copy(fe[iter - 1].begin(), fe[iter - 1].end(), out);
  return 0;
}

void* ptycho_phaser_engine(ptycho_phaser* h) { return h ? (void*)h->ph.eng.ctx : nullptr; }

double ptycho_phaser_phase_seconds(ptycho_phaser* h) { return h ? h->ph.phase_seconds : 0.0; }

}  // extern "C"

namespace {
int mesh_out(const ptycho::ScanMesh& m, float* pos_out, int32_t* canvas_out, float* extent_out) {
  for (size_t q = 0; q < m.positions.size(); ++q) {
    pos_out[2 * q] = m.positions[q].x;
    pos_out[2 * q + 1] = m.positions[q].y;
  }
  if (canvas_out) {
    canvas_out[0] = m.canvas_x;
    canvas_out[1] = m.canvas_y;
  }
  if (extent_out) {
    extent_out[0] = m.minima.x;
    extent_out[1] = m.minima.y;
    extent_out[2] = m.maxima.x;
    extent_out[3] = m.maxima.y;
  }
  return 0;
}
ptycho::Parameters mesh_params(const ptycho_mesh_params* p) {
  ptycho::Parameters prm;
  prm.e.scanDims_x = (unsigned)p->nx;
  prm.e.scanDims_y = (unsigned)p->ny;
  prm.e.stepSize_x = (float)p->step_x;
  prm.e.stepSize_y = (float)p->step_y;
  prm.e.dx_s = p->dx_s;
  prm.e.drift_x = (float)p->drift_x;
  prm.e.drift_y = (float)p->drift_y;
  prm.r.jitterRadius = (unsigned)p->jitter_radius;
  prm.r.flipScanAxis = p->flip != 0;
  prm.r.desiredShape = (unsigned)p->probe_size;
  prm.r.seed = p->seed;
  return prm;
}
}  // namespace

extern "C" {

int ptycho_mesh_list(const ptycho_mesh_params* p, const double* metres, int64_t n, float* pos_out, int32_t* canvas_out,
                     float* extent_out) {
  if (!p || !metres || !pos_out || n <= 0 || !(p->dx_s > 0) || p->probe_size <= 0) return -1;
  ptycho::ListScanMesh m;
  m.metres.assign(metres, metres + 2 * (size_t)n);
  std::string err;
  if (!m.generate(mesh_params(p), &err)) return -1;
  return mesh_out(m, pos_out, canvas_out, extent_out);
}

int ptycho_mesh_spiral(const ptycho_mesh_params* p, float* pos_out, int32_t* canvas_out, float* extent_out) {
  if (!p || !pos_out || p->nx <= 0 || p->ny <= 0 || !(p->dx_s > 0) || p->probe_size <= 0) return -1;
  ptycho::SpiralScanMesh m;
  std::string err;
  if (!m.generate(mesh_params(p), &err)) return -1;
  return mesh_out(m, pos_out, canvas_out, extent_out);
}

int ptycho_mls_probe_modes(void* engine, int32_t m, int32_t k, double lambda, double dx_s, double beam_size,
                           const float* draws, float* out) {
  if (!engine || !out || m < 2 || k < 1 || (k > 1 && !draws) || !(lambda > 0) || !(dx_s > 0)) return -1;
  std::vector<float> modes(out, out + 2 * (size_t)k * m * m);
  std::string err;
  const int rc = ptycho::init_mlh_modes(static_cast<ptyk_ctx*>(engine), m, k, lambda, dx_s, beam_size, draws, true,
                                        &modes, &err);
  if (rc != PTYK_OK) return -1;
  std::copy(modes.begin(), modes.end(), out);
  return 0;
}

int ptycho_save_complex_csv(const char* path, const float* c64, int32_t rows, int32_t cols, int32_t binary) {
  return path && c64 && ptycho::save_complex(path, c64, rows, cols, binary != 0) ? 0 : -1;
}

int ptycho_load_complex_csv(const char* path, float* out, int32_t max_elems, int32_t* rows, int32_t* cols) {
  std::vector<float> v;
  int r = 0, c = 0;
  if (!path || !rows || !cols || !ptycho::load_complex_csv(path, &v, &r, &c)) return -1;
  *rows = r;
  *cols = c;
  if (out) {
    if ((int64_t)r * c > max_elems) return -1;
    std::copy(v.begin(), v.end(), out);
  }
  return 0;
}

}  // extern "C"
