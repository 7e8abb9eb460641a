// C ABI of the engine (include/ptyk.h): handle, device memory, validation, dispatch.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ptyk.h"
#include "ptyk_device.h"
#include "ptyk_internal.h"

using namespace ptyk;

struct ptyk_ctx {
  ptyk_config cfg{};
  hipStream_t stream = nullptr;
  std::string err;
  long mm = 0;  // M*M
  // device buffers
  float2* tw = nullptr;
  float2* obj = nullptr;
  float* obj_int = nullptr;
  float2* probe = nullptr;
  float* probe_int = nullptr;
  float2* probe_tmp = nullptr;
  __half* amp = nullptr;
  float* amp_scale = nullptr;
  float* amp_sum = nullptr;
  float* avg = nullptr;
  float2* pos_dev = nullptr;
  int2* roi_dev = nullptr;
  int* order_dev = nullptr;
  double* scal = nullptr;
  float* factors = nullptr;
  // ePIE scratch
  float2* e_psi_old = nullptr;
  float2* e_psi = nullptr;
  // DM state
  float2* psi = nullptr;
  int* row_ptr = nullptr;
  int* row_frames = nullptr;
  double2* probe_partial = nullptr;
  double* obj_partial = nullptr;
  double* probe_partial_sum = nullptr;
  int probe_chunks = 1, frames_per_chunk = 1;
  // wave-scheduled DM
  int dm_wave_cfg = 0;  // tuning: 0 auto, -1 all frames per pass, k > 0 frames per wave
  int mls_mod_cb = 16, mls_pum_fpc = 4;  // tuning: MLs modulus block width, probe-mode chunk frames
  int mls_row_waves = 1;                   // tuning: wave-tile MLs row passes (M <= 256)
  int mls_graphs = 0;                    // tuning: MLs groups replayed from captured graphs (off: slower here)
  int dm_graph = 0;                      // tuning: DM iterations replayed from captured graphs
  std::vector<std::pair<uint32_t, hipGraphExec_t>> dm_graphs;  // one per flag set (dm_graph)
  std::vector<uintptr_t> dm_graph_sig;   // the schedule they were captured with
  int dm_wave = 0, dm_waves = 0;
  double2* d_obj = nullptr;
  double2* d_probe = nullptr;
  int* wave_row_ptr = nullptr;
  int* wave_row_frames = nullptr;
  std::vector<int> wave_lo, wave_hi;
  // bumped whenever the DM launch tables or shapes may change (positions, tuning, a rebuild):
  // part of the batch graph signature, so a captured graph never replays stale row ranges
  uint64_t dm_gen = 0;
  hipEvent_t ev[8] = {};
  std::vector<hipEvent_t> evpool;
  std::vector<int> evcat;
  // object-error state
  float2* truth = nullptr;
  float* cov = nullptr;
  float* pint = nullptr;
  double* epart = nullptr;
  // MLs state
  MlsDev mls;
  bool mls_ready = false;
  // host state
  std::vector<float2> pos;
  std::vector<int2> roi;
  bool have_frames = false, have_pos = false, have_probe = false;
  bool dm_started = false, modal = false, modes_initialized = false;
};

static thread_local std::string g_create_err;

static int fail(ptyk_ctx* c, int code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    g_create_err = msg;
  return code;
}

#define HIPCHK(c, call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((c), e_ == hipErrorOutOfMemory ? PTYK_ENOMEM : PTYK_EDEVICE,            \
                  std::string(#call) + ": " + hipGetErrorString(e_));                      \
  } while (0)

// Zero-initialised device allocation (Cuda2DArray::allocateDeviceMemory zero-fills,
// Cuda2DArray.hpp:154-160).  The engine stream is non-blocking, so the fill must have
// completed before any later launch on that stream may read the buffer.
template <class T>
static hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e == hipSuccess) e = hipMemset(*p, 0, count * sizeof(T));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return e;
}

static void free_mls(ptyk_ctx* c) {
  MlsDev& d = c->mls;
  if (d.enq) {  // no task may be left to touch the streams and events released below
    (void)d.enq->drain();
    d.enq.reset();
  }
  if (d.side) (void)hipStreamSynchronize(d.side);
  if (d.side2) (void)hipStreamSynchronize(d.side2);
  void* ptrs[] = {d.vmode, d.evo, d.illum, d.pos, d.pos_o, d.roi, d.sub, d.step_obj, d.step_probe, d.ferr, d.flag,
                  d.op, d.vp, d.psi, d.pu, d.spec, d.pum, d.acc, d.precond, d.part, d.gscal, d.gframes,
                  d.ftot, d.dxb, d.pos_part, d.pos_ftot, d.pum_part, d.vp_part, d.vp_ftot, d.op_alt, d.dxb_alt,
                  d.goffs_dev, d.gsel};  // tmp, tmp2 in psi
  mls_drop_graphs(d);
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (hipEvent_t ev : d.ev_side)
    if (ev) hipEventDestroy(ev);
  for (hipEvent_t ev : d.ev_join)
    if (ev) hipEventDestroy(ev);
  if (d.ev_vpdone) hipEventDestroy(d.ev_vpdone);
  if (d.side) hipStreamDestroy(d.side);
  if (d.side2) hipStreamDestroy(d.side2);
  d = MlsDev();
  c->mls_ready = false;
}

static void drop_dm_graphs(ptyk_ctx* c) {
  for (auto& g : c->dm_graphs) hipGraphExecDestroy(g.second);
  c->dm_graphs.clear();
}

static void free_all(ptyk_ctx* c) {
  drop_dm_graphs(c);
  free_mls(c);
  for (void* p : {(void*)c->truth, (void*)c->cov, (void*)c->pint, (void*)c->epart})
    if (p) hipFree(p);
  void* ptrs[] = {c->tw, c->obj, c->obj_int, c->probe, c->probe_int, c->probe_tmp, c->amp, c->amp_scale,
                  c->amp_sum, c->avg, c->pos_dev, c->roi_dev, c->order_dev, c->scal, c->factors, c->e_psi_old,
                  c->e_psi, c->psi, c->row_ptr, c->row_frames, c->probe_partial, c->obj_partial,
                  c->probe_partial_sum, c->d_obj, c->d_probe, c->wave_row_ptr, c->wave_row_frames};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (auto& e : c->ev)
    if (e) hipEventDestroy(e);
  for (auto& e : c->evpool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
}

extern "C" {

int ptyk_create(const ptyk_config* cfg, ptyk_ctx** out) {
  if (!cfg || !out) return fail(nullptr, PTYK_EINVAL, "null argument");
  *out = nullptr;
  if (!frame_size_supported(cfg->frame_size))
    return fail(nullptr, PTYK_EINVAL, "frame_size must be one of 16, 32, 64, 128, 256, 512");
  if (cfg->n_frames <= 0 || cfg->n_modes <= 0 || cfg->n_modes > 16 || cfg->obj_rows < cfg->frame_size ||
      cfg->obj_cols < cfg->frame_size)
    return fail(nullptr, PTYK_EINVAL, "bad n_frames / n_modes / object shape");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, PTYK_EDEVICE, "no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, PTYK_EINVAL, "bad device ordinal");
  if (hipSetDevice(cfg->device) != hipSuccess) return fail(nullptr, PTYK_EDEVICE, "hipSetDevice failed");

  ptyk_ctx* c = new ptyk_ctx();
  c->cfg = *cfg;
  const long n = cfg->n_frames, m = cfg->frame_size, k = cfg->n_modes;
  c->mm = m * m;
  const long canvas = (long)cfg->obj_rows * cfg->obj_cols;
  auto bail = [&](hipError_t e, const char* what) {
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    free_all(c);
    delete c;
    return fail(nullptr, e == hipErrorOutOfMemory ? PTYK_ENOMEM : PTYK_EDEVICE, msg);
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
  if ((e = noalloc_setup(cfg->device)) != hipSuccess) return bail(e, "scratch array address");
  if ((e = dalloc(&c->tw, PTYK_TWN)) != hipSuccess) return bail(e, "twiddles");
  {
    std::vector<float2> tw(PTYK_TWN);
    for (int i = 0; i < PTYK_TWN; ++i) {
      const double a = -2.0 * M_PI * (double)i / PTYK_TWN;
      tw[i] = make_float2((float)cos(a), (float)sin(a));
    }
    if ((e = hipMemcpy(c->tw, tw.data(), PTYK_TWN * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess)
      return bail(e, "twiddle upload");
  }
  if ((e = dalloc(&c->obj, canvas)) != hipSuccess) return bail(e, "object");
  if ((e = dalloc(&c->obj_int, canvas)) != hipSuccess) return bail(e, "object intensities");
  if ((e = dalloc(&c->probe, k * c->mm)) != hipSuccess) return bail(e, "probe");
  if ((e = dalloc(&c->probe_int, k * c->mm)) != hipSuccess) return bail(e, "probe intensities");
  if ((e = dalloc(&c->probe_tmp, k * c->mm)) != hipSuccess) return bail(e, "probe scratch");
  if ((e = dalloc(&c->amp, n * c->mm)) != hipSuccess) return bail(e, "frames");
  if ((e = dalloc(&c->amp_scale, n)) != hipSuccess) return bail(e, "frame scales");
  if ((e = dalloc(&c->amp_sum, n)) != hipSuccess) return bail(e, "frame sums");
  if ((e = dalloc(&c->avg, c->mm)) != hipSuccess) return bail(e, "mean frame");
  if ((e = dalloc(&c->pos_dev, n)) != hipSuccess) return bail(e, "positions");
  if ((e = dalloc(&c->roi_dev, n)) != hipSuccess) return bail(e, "roi");
  if ((e = dalloc(&c->order_dev, n)) != hipSuccess) return bail(e, "order");
  if ((e = dalloc(&c->scal, PTYK_SCAL_COUNT)) != hipSuccess) return bail(e, "scalars");
  if ((e = dalloc(&c->factors, 2)) != hipSuccess) return bail(e, "factors");
  if ((e = dalloc(&c->e_psi_old, k * c->mm)) != hipSuccess) return bail(e, "epie scratch");
  if ((e = dalloc(&c->e_psi, k * c->mm)) != hipSuccess) return bail(e, "epie scratch");
  for (auto& ev : c->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(e, "events");
  // Probe::m_maxIntensity starts at 1 (Probe.cpp:54-58); objMax at 0.
  {
    double s0[PTYK_SCAL_COUNT] = {0.0, 1.0};
    if ((e = hipMemcpy(c->scal, s0, sizeof(s0), hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "scalars");
  }
  *out = c;
  return PTYK_OK;
}

void ptyk_destroy(ptyk_ctx* c) {
  if (!c) return;
  hipSetDevice(c->cfg.device);
  hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

const char* ptyk_last_error(const ptyk_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int ptyk_synchronize(ptyk_ctx* c) {
  if (!c) return PTYK_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

// ---------------------------------------------------------------- frames
int ptyk_set_frames_f32(ptyk_ctx* c, const float* a) {
  if (!c || !a) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  const long n = c->cfg.n_frames, total = n * c->mm;
  for (long i = 0; i < total; ++i)
    if (!std::isfinite(a[i]) || a[i] < 0.f) return fail(c, PTYK_EINVAL, "amplitudes must be finite and >= 0");
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, total * sizeof(float)));
  hipError_t e = hipMemcpyAsync(tmp, a, total * sizeof(float), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = quantize_frames(tmp, c->amp, c->amp_scale, n, c->cfg.frame_size, c->stream);
  if (e == hipSuccess) e = frame_sums(c->amp, c->amp_scale, c->amp_sum, n, c->cfg.frame_size, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  HIPCHK(c, e);
  c->have_frames = true;
  return PTYK_OK;
}

int ptyk_set_frames_f16(ptyk_ctx* c, const uint16_t* h, const float* sc) {
  if (!c || !h || !sc) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  const long n = c->cfg.n_frames;
  for (long i = 0; i < n; ++i)
    if (!(sc[i] > 0.f) || !std::isfinite(sc[i])) return fail(c, PTYK_EINVAL, "scales must be finite and > 0");
  HIPCHK(c, hipMemcpyAsync(c->amp, h, n * c->mm * sizeof(__half), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->amp_scale, sc, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, frame_sums(c->amp, c->amp_scale, c->amp_sum, n, c->cfg.frame_size, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_frames = true;
  return PTYK_OK;
}

int ptyk_get_frames_f32(ptyk_ctx* c, float* out) {
  if (!c || !out) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->have_frames) return fail(c, PTYK_ESTATE, "frames not set");
  hipSetDevice(c->cfg.device);
  const long n = c->cfg.n_frames, total = n * c->mm;
  float* tmp = nullptr;
  HIPCHK(c, hipMalloc((void**)&tmp, total * sizeof(float)));
  hipError_t e = dequantize_frames(c->amp, c->amp_scale, tmp, n, c->cfg.frame_size, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, total * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  HIPCHK(c, e);
  return PTYK_OK;
}

int ptyk_get_frame_scales(ptyk_ctx* c, float* out) {
  if (!c || !out) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  HIPCHK(c, hipMemcpyAsync(out, c->amp_scale, c->cfg.n_frames * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

// ---------------------------------------------------------------- positions
static void free_dm(ptyk_ctx* c) {
  void* ptrs[] = {c->psi, c->row_ptr, c->row_frames, c->probe_partial, c->obj_partial, c->probe_partial_sum,
                  c->d_obj, c->d_probe, c->wave_row_ptr, c->wave_row_frames};
  for (void* p : ptrs)
    if (p) hipFree(p);
  c->psi = nullptr;
  c->row_ptr = c->row_frames = nullptr;
  c->probe_partial = nullptr;
  c->obj_partial = c->probe_partial_sum = nullptr;
  c->d_obj = c->d_probe = nullptr;
  c->wave_row_ptr = c->wave_row_frames = nullptr;
  c->dm_wave = c->dm_waves = 0;
  c->dm_started = false;
  ++c->dm_gen;
}

int ptyk_set_positions(ptyk_ctx* c, const float* p) {
  if (!c || !p) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  const int n = c->cfg.n_frames, m = c->cfg.frame_size;
  std::vector<float2> pos(n);
  std::vector<int2> roi(n);
  for (int i = 0; i < n; ++i) {
    const float x = p[2 * i], y = p[2 * i + 1];
    if (!std::isfinite(x) || !std::isfinite(y)) return fail(c, PTYK_EINVAL, "non-finite position");
    // ePIE.cpp:100-105 rejects negative positions; ROIs must also fit in the canvas.
    if (x < 0.f || y < 0.f)
      return fail(c, PTYK_ERANGE, "position " + std::to_string(i) + " is negative: object array too small");
    const int rx = (int)roundf(x), ry = (int)roundf(y);
    if (rx + m > c->cfg.obj_rows || ry + m > c->cfg.obj_cols)
      return fail(c, PTYK_ERANGE, "ROI of position " + std::to_string(i) + " exceeds the object canvas");
    pos[i] = make_float2(x, y);
    roi[i] = make_int2(rx, ry);
  }
  HIPCHK(c, hipMemcpyAsync(c->pos_dev, pos.data(), n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->roi_dev, roi.data(), n * sizeof(int2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->pos = pos;
  c->roi = roi;
  c->have_pos = true;
  ++c->dm_gen;
  // DM's frame lists depend on the positions; rebuild them on the next DM call.
  if (c->row_ptr) {
    hipFree(c->row_ptr);
    hipFree(c->row_frames);
    c->row_ptr = c->row_frames = nullptr;
  }
  if (c->wave_row_ptr) {
    hipFree(c->wave_row_ptr);
    hipFree(c->wave_row_frames);
    c->wave_row_ptr = c->wave_row_frames = nullptr;
    c->dm_wave = c->dm_waves = 0;
  }
  return PTYK_OK;
}

int ptyk_get_roi(ptyk_ctx* c, int32_t* out) {
  if (!c || !out) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->have_pos) return fail(c, PTYK_ESTATE, "positions not set");
  for (size_t i = 0; i < c->roi.size(); ++i) {
    out[2 * i] = c->roi[i].x;
    out[2 * i + 1] = c->roi[i].y;
  }
  return PTYK_OK;
}

// ---------------------------------------------------------------- arrays
static int copy_in(ptyk_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !src) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}
static int copy_out(ptyk_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

int ptyk_set_object(ptyk_ctx* c, const float* o) {
  if (!c) return PTYK_EINVAL;
  const long canvas = (long)c->cfg.obj_rows * c->cfg.obj_cols;
  int r = copy_in(c, c->obj, o, canvas * sizeof(float2));
  if (r) return r;
  HIPCHK(c, intensities(c->obj, c->obj_int, canvas, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}
int ptyk_get_object(ptyk_ctx* c, float* o) {
  if (!c) return PTYK_EINVAL;
  return copy_out(c, o, c->obj, (size_t)c->cfg.obj_rows * c->cfg.obj_cols * sizeof(float2));
}
int ptyk_clear_object(ptyk_ctx* c) {
  if (!c) return PTYK_EINVAL;
  hipSetDevice(c->cfg.device);
  const long canvas = (long)c->cfg.obj_rows * c->cfg.obj_cols;
  HIPCHK(c, hipMemsetAsync(c->obj, 0, canvas * sizeof(float2), c->stream));
  HIPCHK(c, hipMemsetAsync(c->obj_int, 0, canvas * sizeof(float), c->stream));
  double z = 0.0;
  HIPCHK(c, hipMemcpyAsync(c->scal + PTYK_SCAL_OBJ_MAX, &z, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}
int ptyk_set_probe(ptyk_ctx* c, const float* p) {
  if (!c) return PTYK_EINVAL;
  const long cnt = (long)c->cfg.n_modes * c->mm;
  int r = copy_in(c, c->probe, p, cnt * sizeof(float2));
  if (r) return r;
  HIPCHK(c, intensities(c->probe, c->probe_int, cnt, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_probe = true;
  return PTYK_OK;
}
int ptyk_get_probe(ptyk_ctx* c, float* p) {
  if (!c) return PTYK_EINVAL;
  return copy_out(c, p, c->probe, (size_t)c->cfg.n_modes * c->mm * sizeof(float2));
}

static int ensure_psi(ptyk_ctx* c) {
  if (c->psi) return PTYK_OK;
  hipError_t e = dalloc(&c->psi, (size_t)c->cfg.n_frames * c->mm);
  if (e != hipSuccess) return fail(c, PTYK_ENOMEM, std::string("psi: ") + hipGetErrorString(e));
  return PTYK_OK;
}
int ptyk_set_psi(ptyk_ctx* c, const float* p) {
  if (!c) return PTYK_EINVAL;
  hipSetDevice(c->cfg.device);
  int r = ensure_psi(c);
  if (r) return r;
  return copy_in(c, c->psi, p, (size_t)c->cfg.n_frames * c->mm * sizeof(float2));
}
int ptyk_get_psi(ptyk_ctx* c, float* p) {
  if (!c) return PTYK_EINVAL;
  if (!c->psi) return fail(c, PTYK_ESTATE, "no DM state yet");
  return copy_out(c, p, c->psi, (size_t)c->cfg.n_frames * c->mm * sizeof(float2));
}

int ptyk_get_scalars(ptyk_ctx* c, double* out) {
  if (!c || !out) return fail(c, PTYK_EINVAL, "null argument");
  int r = copy_out(c, out, c->scal, PTYK_SCAL_COUNT * sizeof(double));
  if (r) return r;
  float f[2];
  r = copy_out(c, f, c->factors, sizeof(f));
  if (r) return r;
  out[PTYK_SCAL_DM_FACTOR_OBJ] = f[0];
  out[PTYK_SCAL_DM_FACTOR_PROBE] = f[1];
  return PTYK_OK;
}
int ptyk_set_scalars(ptyk_ctx* c, const double* in) {
  if (!c || !in) return fail(c, PTYK_EINVAL, "null argument");
  int r = copy_in(c, c->scal, in, PTYK_SCAL_COUNT * sizeof(double));
  if (r) return r;
  const float f[2] = {(float)in[PTYK_SCAL_DM_FACTOR_OBJ], (float)in[PTYK_SCAL_DM_FACTOR_PROBE]};
  r = copy_in(c, c->factors, f, sizeof(f));
  if (r) return r;
  c->dm_started = true;
  return PTYK_OK;
}

int ptyk_refresh_maxima(ptyk_ctx* c) {
  if (!c) return PTYK_EINVAL;
  hipSetDevice(c->cfg.device);
  const long canvas = (long)c->cfg.obj_rows * c->cfg.obj_cols;
  HIPCHK(c, intensities(c->obj, c->obj_int, canvas, c->stream));
  HIPCHK(c, max_abs2(c->obj, canvas, c->scal + PTYK_SCAL_OBJ_MAX, c->stream));
  HIPCHK(c, intensities(c->probe, c->probe_int, (long)c->cfg.n_modes * c->mm, c->stream));
  HIPCHK(c, max_abs2(c->probe, (long)c->cfg.n_modes * c->mm, c->scal + PTYK_SCAL_PROBE_MAX, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

// ---------------------------------------------------------------- probe init
int ptyk_init_probe(ptyk_ctx* c, double beam_size, double dx_s) {
  if (!c) return PTYK_EINVAL;
  if (!c->have_frames) return fail(c, PTYK_ESTATE, "frames not set");
  if (!(beam_size > 0) || !(dx_s > 0)) return fail(c, PTYK_EINVAL, "beam_size and dx_s must be > 0");
  hipSetDevice(c->cfg.device);
  const int m = c->cfg.frame_size;
  const long cnt = (long)c->cfg.n_modes * c->mm;
  // h_initProbe(..., gaussWindow = M/(2*2.35), beamSize, dx_s) with real_t (float) arguments
  const float window = (float)((double)m / (2 * 2.35));
  const float factor = (float)dx_s / (float)beam_size;
  HIPCHK(c, hipMemsetAsync(c->probe, 0, cnt * sizeof(float2), c->stream));
  HIPCHK(c, probe_sinc_gauss(c->probe, m, window, factor, c->stream));
  HIPCHK(c, mean_amplitude(c->amp, c->amp_scale, c->avg, c->cfg.n_frames, m, c->stream));
  HIPCHK(c, fft2_batch(c->probe, 1, m, false, c->tw, c->stream));
  HIPCHK(c, modulus_project_single(c->probe, c->avg, m, c->stream));
  HIPCHK(c, fft2_batch(c->probe, 1, m, true, c->tw, c->stream));
  HIPCHK(c, probe_normalize(c->probe, c->probe_int, c->cfg.n_modes, m, c->scal + PTYK_SCAL_PROBE_MAX, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_probe = true;
  c->modes_initialized = false;
  return PTYK_OK;
}

int ptyk_begin_modal(ptyk_ctx* c) {
  if (!c) return PTYK_EINVAL;
  hipSetDevice(c->cfg.device);
  c->modal = true;
  if (c->modes_initialized) return PTYK_OK;
  const int k = c->cfg.n_modes;
  if (k > 1) {
    HIPCHK(c, modes_from_base(c->probe, c->probe_int, k, c->cfg.frame_size, 0.05f, c->stream));
    HIPCHK(c, orthogonalize_modes(c->probe, c->probe_tmp, k, c->cfg.frame_size, c->stream));
    c->modes_initialized = true;  // Probe::initProbeModes sets it only when K > 1
  }
  // initProbeModes ends with updateMaxIntensity(): the modal maximum of the stored
  // intensities, which orthogonalize() leaves untouched (Probe.cpp:320-330).
  const long mm = c->mm;
  std::vector<float> pint((size_t)k * mm);
  HIPCHK(c, hipMemcpyAsync(pint.data(), c->probe_int, pint.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float mx = 0.f;
  for (long q = 0; q < mm; ++q) {
    float s = 0.f;
    for (int i = 0; i < k; ++i) s += pint[(size_t)i * mm + q];
    mx = std::max(mx, s);
  }
  const double v = mx;
  HIPCHK(c, hipMemcpyAsync(c->scal + PTYK_SCAL_PROBE_MAX, &v, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

int ptyk_orthogonalize_modes(ptyk_ctx* c) {
  if (!c) return PTYK_EINVAL;
  if (!c->have_probe) return fail(c, PTYK_ESTATE, "probe must be set");
  hipSetDevice(c->cfg.device);
  if (c->cfg.n_modes > 1) {
    HIPCHK(c, orthogonalize_modes(c->probe, c->probe_tmp, c->cfg.n_modes, c->cfg.frame_size, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return PTYK_OK;
}

int ptyk_end_modal(ptyk_ctx* c) {
  if (!c) return PTYK_EINVAL;
  c->modal = false;
  c->modes_initialized = false;
  return PTYK_OK;
}

// ---------------------------------------------------------------- ePIE
int ptyk_epie_iteration(ptyk_ctx* c, const int32_t* order, uint32_t flags, double* err) {
  if (!c || !order) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->have_frames || !c->have_pos || !c->have_probe)
    return fail(c, PTYK_ESTATE, "frames, positions and probe must be set");
  hipSetDevice(c->cfg.device);
  const int n = c->cfg.n_frames;
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    if (order[i] < 0 || order[i] >= n || seen[order[i]]) return fail(c, PTYK_EINVAL, "order must be a permutation");
    seen[order[i]] = 1;
  }
  if (flags & PTYK_FLAG_UPDATE_MODES) {
    int r = ptyk_begin_modal(c);
    if (r) return r;
  }
  HIPCHK(c, hipMemcpyAsync(c->order_dev, order, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  EpieArgs a{};
  a.obj = c->obj;
  a.obj_int = c->obj_int;
  a.probe = c->probe;
  a.probe_int = c->probe_int;
  a.psi_old = c->e_psi_old;
  a.psi = c->e_psi;
  a.amp = c->amp;
  a.amp_scale = c->amp_scale;
  a.amp_sum = c->amp_sum;
  a.pos = c->pos_dev;
  a.order = c->order_dev;
  a.scal = c->scal;
  a.tw = c->tw;
  a.n_pos = n;
  a.m = c->cfg.frame_size;
  a.modal = c->modal && c->cfg.n_modes > 1;
  a.k_use = a.modal ? c->cfg.n_modes : 1;
  a.k_total = c->cfg.n_modes;
  a.obj_rows = c->cfg.obj_rows;
  a.obj_cols = c->cfg.obj_cols;
  a.phase_constraint = (flags & PTYK_FLAG_PHASE_CONSTRAINT) ? 1 : 0;
  a.update_probe = (flags & PTYK_FLAG_UPDATE_PROBE) ? 1 : 0;
  a.rms = (flags & PTYK_FLAG_RMS) ? 1 : 0;
  a.tex_quantize = c->cfg.tex_quantize;
  HIPCHK(c, epie_iteration(a, c->stream));
  double s[PTYK_SCAL_COUNT];
  HIPCHK(c, hipMemcpyAsync(s, c->scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (err) *err = a.rms ? s[PTYK_SCAL_ERROR] / n : 0.0;
  return PTYK_OK;
}

// ---------------------------------------------------------------- DM
static int dm_wave_frames(const ptyk_ctx* c) {
  const int n = c->cfg.n_frames, m = c->cfg.frame_size;
  if (c->dm_wave_cfg < 0 || !dm_wave_supported(m)) return 0;  // one pass over all frames
  if (c->dm_wave_cfg > 0) return std::min(c->dm_wave_cfg, n);
  // a wave's in-place intermediates (W x M x M complex64) kept near 128 MiB (measured best of
  // 16..128 frames of 512^2 on MI355X)
  const long w = (128L << 20) / ((long)m * m * 8);
  return (int)std::max(1L, std::min((long)n, w));
}

static int dm_prepare(ptyk_ctx* c) {
  const int n = c->cfg.n_frames, m = c->cfg.frame_size, rows = c->cfg.obj_rows;
  int r = ensure_psi(c);
  if (r) return r;
  const int wave = dm_wave_frames(c);
  if (wave == 0 && !c->row_ptr) {
    // frames covering each object row, ascending frame id
    std::vector<int> cnt(rows + 1, 0);
    for (int j = 0; j < n; ++j)
      for (int q = 0; q < m; ++q) cnt[c->roi[j].x + q + 1]++;
    for (int p = 0; p < rows; ++p) cnt[p + 1] += cnt[p];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1), frames((size_t)n * m);
    for (int j = 0; j < n; ++j)
      for (int q = 0; q < m; ++q) frames[fill[c->roi[j].x + q]++] = j;
    HIPCHK(c, dalloc(&c->row_ptr, rows + 1));
    HIPCHK(c, dalloc(&c->row_frames, (size_t)n * m));
    HIPCHK(c, hipMemcpyAsync(c->row_ptr, cnt.data(), (rows + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->row_frames, frames.data(), frames.size() * sizeof(int), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors die at the end of this block
    ++c->dm_gen;
  }
  if (wave > 0 && c->dm_wave != wave) {
    // per wave: frames covering each object row (offsets into one frame-id array)
    const int nw = (n + wave - 1) / wave;
    std::vector<int> ptr((size_t)nw * (rows + 1)), frames((size_t)n * m);
    std::vector<int> lo(nw), hi(nw);
    int off = 0;
    for (int w = 0; w < nw; ++w) {
      const int f0 = w * wave, f1 = std::min(n, f0 + wave);
      std::vector<int> cnt(rows + 1, 0);
      int rlo = rows, rhi = 0;
      for (int j = f0; j < f1; ++j) {
        rlo = std::min(rlo, c->roi[j].x);
        rhi = std::max(rhi, c->roi[j].x + m);
        for (int q = 0; q < m; ++q) cnt[c->roi[j].x + q + 1]++;
      }
      for (int p = 0; p < rows; ++p) cnt[p + 1] += cnt[p];
      int* wp = ptr.data() + (size_t)w * (rows + 1);
      for (int p = 0; p <= rows; ++p) wp[p] = off + cnt[p];
      std::vector<int> fill(rows);
      for (int p = 0; p < rows; ++p) fill[p] = wp[p];
      for (int j = f0; j < f1; ++j)
        for (int q = 0; q < m; ++q) frames[fill[c->roi[j].x + q]++] = j;
      off += cnt[rows];
      lo[w] = rlo;
      hi[w] = rhi;
    }
    if (c->wave_row_ptr) hipFree(c->wave_row_ptr);
    if (c->wave_row_frames) hipFree(c->wave_row_frames);
    c->wave_row_ptr = c->wave_row_frames = nullptr;
    HIPCHK(c, dalloc(&c->wave_row_ptr, ptr.size()));
    HIPCHK(c, dalloc(&c->wave_row_frames, frames.size()));
    HIPCHK(c, hipMemcpyAsync(c->wave_row_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->wave_row_frames, frames.data(), frames.size() * sizeof(int), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors die at the end of this block
    c->wave_lo = lo;
    c->wave_hi = hi;
    c->dm_wave = wave;
    c->dm_waves = nw;
    ++c->dm_gen;
    if (!c->d_obj) HIPCHK(c, dalloc(&c->d_obj, (size_t)rows * c->cfg.obj_cols));
    if (!c->d_probe) HIPCHK(c, dalloc(&c->d_probe, (size_t)c->mm));
  }
  if (!c->probe_partial) {
    const int rpw = m >= PTYK_NT ? 1 : PTYK_NT / m;
    const int gp = m / rpw;
    c->probe_chunks = std::max(1, std::min(n, 2048 / gp));
    c->frames_per_chunk = (n + c->probe_chunks - 1) / c->probe_chunks;
    c->probe_chunks = (n + c->frames_per_chunk - 1) / c->frames_per_chunk;
    HIPCHK(c, dalloc(&c->probe_partial, (size_t)c->probe_chunks * c->mm));
    HIPCHK(c, dalloc(&c->obj_partial, (size_t)rows * dm_obj_col_chunks(c->cfg.obj_cols)));
    HIPCHK(c, dalloc(&c->probe_partial_sum, (size_t)std::max(gp, (int)((c->mm + PTYK_NT - 1) / PTYK_NT))));
  }
  if (!c->dm_started) {
    // DM.cu:88-92: cached factors {1/objMax, 1/probeMax} from the current maxima
    double s[PTYK_SCAL_COUNT];
    HIPCHK(c, hipMemcpyAsync(s, c->scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const float f[2] = {(float)(1.0 / (double)(float)s[0]), (float)(1.0 / (double)(float)s[1])};
    HIPCHK(c, hipMemcpyAsync(c->factors, f, sizeof(f), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // f lives on this stack frame
    c->dm_started = true;
  }
  return PTYK_OK;
}

// The launches of one DM iteration on the context's stream (no host synchronisation, no
// allocation: safe inside a stream capture once dm_prepare has run).
// HIP events (and their category slots) of the launch timers, grown on demand.
static int timer_reserve(ptyk_ctx* c, size_t n) {
  while (c->evpool.size() < n) {
    hipEvent_t e = nullptr;
    HIPCHK(c, hipEventCreate(&e));
    c->evpool.push_back(e);
  }
  c->evcat.resize(c->evpool.size());
  return PTYK_OK;
}

static int dm_launch(ptyk_ctx* c, uint32_t flags, float* kernel_ms) {
  DmTimer tm;
  if (kernel_ms) {
    int r = timer_reserve(c, 1024);
    if (r) return r;
    tm.ev = c->evpool.data();
    tm.cat = c->evcat.data();
    tm.cap = (int)c->evpool.size();
  }
  const int pc = (flags & PTYK_FLAG_PHASE_CONSTRAINT) ? 1 : 0;
  const int up = (flags & PTYK_FLAG_UPDATE_PROBE) ? 1 : 0;
  if (dm_wave_frames(c) > 0) {
    DmWaveArgs a{};
    a.psi = c->psi;
    a.obj = c->obj;
    a.probe = c->probe;
    a.amp = c->amp;
    a.amp_scale = c->amp_scale;
    a.roi = c->roi_dev;
    a.wave_row_ptr = c->wave_row_ptr;
    a.wave_row_frames = c->wave_row_frames;
    a.wave_row_lo = c->wave_lo.data();
    a.wave_row_hi = c->wave_hi.data();
    a.d_obj = c->d_obj;
    a.d_probe = c->d_probe;
    a.obj_partial = c->obj_partial;
    a.probe_partial_sum = c->probe_partial_sum;
    a.factors = c->factors;
    a.scal = c->scal;
    a.tw = c->tw;
    a.n_frames = c->cfg.n_frames;
    a.m = c->cfg.frame_size;
    a.obj_rows = c->cfg.obj_rows;
    a.obj_cols = c->cfg.obj_cols;
    a.wave = c->dm_wave;
    a.n_waves = c->dm_waves;
    a.phase_constraint = pc;
    a.update_probe = up;
    HIPCHK(c, dm_wave_iteration(a, c->stream, kernel_ms ? &tm : nullptr));
  } else {
    DmArgs a{};
    a.psi = c->psi;
    a.obj = c->obj;
    a.probe = c->probe;
    a.amp = c->amp;
    a.amp_scale = c->amp_scale;
    a.roi = c->roi_dev;
    a.row_ptr = c->row_ptr;
    a.row_frames = c->row_frames;
    a.probe_partial = c->probe_partial;
    a.obj_partial = c->obj_partial;
    a.probe_partial_sum = c->probe_partial_sum;
    a.factors = c->factors;
    a.scal = c->scal;
    a.tw = c->tw;
    a.n_frames = c->cfg.n_frames;
    a.m = c->cfg.frame_size;
    a.obj_rows = c->cfg.obj_rows;
    a.obj_cols = c->cfg.obj_cols;
    a.probe_chunks = c->probe_chunks;
    a.frames_per_chunk = c->frames_per_chunk;
    a.phase_constraint = pc;
    a.update_probe = up;
    HIPCHK(c, dm_iteration(a, c->stream, kernel_ms ? &tm : nullptr));
  }
  if (kernel_ms) {
    for (int i = 0; i < DM_T_COUNT; ++i) kernel_ms[i] = 0.f;
    if (tm.n > 0) HIPCHK(c, hipEventSynchronize(tm.ev[tm.n - 1]));
    for (int i = 0; i + 1 < tm.n; ++i) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, tm.ev[i], tm.ev[i + 1]));
      if (tm.cat[i] >= 0 && tm.cat[i] < DM_T_COUNT) kernel_ms[tm.cat[i]] += ms;
    }
  }
  return PTYK_OK;
}

static std::vector<uintptr_t> dm_signature(const ptyk_ctx* c);

// With the DM-graph knob, an iteration after the first replays a graph captured once per flag set
// on the handle's stream (re-captured when the schedule's buffers or frame lists change: the
// signature holds their addresses and the table generation counter).
static int dm_run(ptyk_ctx* c, uint32_t flags, float* kernel_ms) {
  if (!c) return PTYK_EINVAL;
  if (!c->have_frames || !c->have_pos || !c->have_probe)
    return fail(c, PTYK_ESTATE, "frames, positions and probe must be set");
  hipSetDevice(c->cfg.device);
  const bool started = c->dm_started && c->psi;  // the first call sets host-side state eagerly
  int r = dm_prepare(c);
  if (r) return r;
  if (!c->dm_graph || kernel_ms || !started) return dm_launch(c, flags, kernel_ms);
  const std::vector<uintptr_t> sig = dm_signature(c);
  if (sig != c->dm_graph_sig) {
    drop_dm_graphs(c);
    c->dm_graph_sig = sig;
  }
  hipGraphExec_t exec = nullptr;
  for (auto& g : c->dm_graphs)
    if (g.first == flags) exec = g.second;
  if (!exec) {
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    r = dm_launch(c, flags, nullptr);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (r != PTYK_OK || ec != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      return r != PTYK_OK ? r : fail(c, PTYK_EDEVICE, "DM graph capture failed");
    }
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(c, PTYK_EDEVICE, std::string("DM graph instantiate: ") + hipGetErrorString(ei));
    c->dm_graphs.emplace_back(flags, exec);
  }
  HIPCHK(c, hipGraphLaunch(exec, c->stream));
  return PTYK_OK;
}

int ptyk_dm_iteration(ptyk_ctx* c, uint32_t flags) { return dm_run(c, flags, nullptr); }
int ptyk_dm_iteration_timed(ptyk_ctx* c, uint32_t flags, float* kernel_ms) {
  if (!kernel_ms) return fail(c, PTYK_EINVAL, "null argument");
  return dm_run(c, flags, kernel_ms);
}

int ptyk_set_tuning(ptyk_ctx* c, int32_t key, int32_t value) {
  if (!c) return PTYK_EINVAL;
  if (key == PTYK_TUNE_DM_WAVE) {
    if (value < -1) return fail(c, PTYK_EINVAL, "dm wave must be -1, 0 or a frame count");
    c->dm_wave_cfg = value;
    ++c->dm_gen;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_MLS_MOD_CB) {
    if (value != 8 && value != 16 && value != 32) return fail(c, PTYK_EINVAL, "MLs modulus block width: 8, 16 or 32");
    c->mls_mod_cb = value;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_MLS_PUM_FPC) {
    if (value < MLS_PUM_FPC_MIN || value > 64)
      return fail(c, PTYK_EINVAL, "MLs probe-mode chunk: 2 to 64 frames");
    c->mls_pum_fpc = value;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_MLS_GRAPHS) {
    if (value != 0 && value != 1) return fail(c, PTYK_EINVAL, "MLs group graphs: 0 or 1");
    c->mls_graphs = value;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_MLS_ROW_WAVES) {
    if (value != 0 && value != 1) return fail(c, PTYK_EINVAL, "MLs row waves: 0 or 1");
    c->mls_row_waves = value;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_DM_GRAPH) {
    if (value != 0 && value != 1) return fail(c, PTYK_EINVAL, "DM graphs: 0 or 1");
    c->dm_graph = value;
    if (!value) drop_dm_graphs(c);
    return PTYK_OK;
  }
  return fail(c, PTYK_EINVAL, "unknown tuning key");
}

int ptyk_get_tuning(ptyk_ctx* c, int32_t key, int32_t* value) {
  if (!c || !value) return PTYK_EINVAL;
  if (key == PTYK_TUNE_DM_WAVE) {
    *value = dm_wave_frames(c);
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_DM_GRAPH) {
    *value = c->dm_graph;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_MLS_ROW_WAVES) {
    *value = c->mls_row_waves;
    return PTYK_OK;
  }
  if (key == PTYK_TUNE_MLS_MOD_CB || key == PTYK_TUNE_MLS_PUM_FPC || key == PTYK_TUNE_MLS_GRAPHS) {
    *value = key == PTYK_TUNE_MLS_MOD_CB ? c->mls_mod_cb : key == PTYK_TUNE_MLS_PUM_FPC ? c->mls_pum_fpc : c->mls_graphs;
    return PTYK_OK;
  }
  return fail(c, PTYK_EINVAL, "unknown tuning key");
}

// ---------------------------------------------------------------- MLs
// Group scratch for groups of up to g frames (grown on demand).
static int mls_reserve(ptyk_ctx* c, int g) {
  MlsDev& d = c->mls;
  const long mm = c->mm;
  const int ch = mls_chunks(d.m);
  const long nb = (mm + PTYK_NT - 1) / PTYK_NT;
  if (!d.side) {
    HIPCHK(c, hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking));
    // fork events: timing-enabled, since a launch's own dispatch may record one (hipExtLaunchKernelGGL).
    // The fork, join and variable-mode events order streams of this one device only: each kernel's
    // own end-of-kernel release (device scope) makes its writes visible to the other streams, and
    // the host reads results through copies queued after a synchronisation, so these events skip
    // the system-scope fences (an L2 write-back and invalidate on every record and wait)
    for (hipEvent_t& ev : d.ev_side) HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
    HIPCHK(c, hipStreamCreateWithFlags(&d.side2, hipStreamNonBlocking));
    for (hipEvent_t& ev : d.ev_join)
      HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming | hipEventDisableSystemFence));
    HIPCHK(c, hipEventCreateWithFlags(&d.ev_vpdone, hipEventDisableTiming | hipEventDisableSystemFence));
  }
  if (!d.enq) d.enq.reset(new SideEnqueuer(c->cfg.device));
  if (g > d.gmax) {
    mls_drop_graphs(d);  // they hold the addresses of the scratch about to move
    void* ptrs[] = {d.op,  d.vp,       d.psi,      d.pu,       d.gscal,   d.ftot,  d.gsel,    d.op_alt,
                    d.dxb, d.pos_part, d.pos_ftot, d.pum_part, d.vp_part, d.vp_ftot, d.dxb_alt};
    for (void* p : ptrs)  // tmp and tmp2 live inside psi
      if (p) hipFree(p);
    d.op = d.vp = d.psi = d.pu = d.tmp = d.tmp2 = d.dxb = d.pum_part = d.op_alt = d.dxb_alt = nullptr;
    d.gsel = nullptr;
    d.gscal = d.ftot = d.pos_part = d.pos_ftot = d.vp_part = d.vp_ftot = nullptr;
    d.gmax = 0;
    const long gm = (long)g * mm;
    const long tcap = std::max<long>(g, d.k);
    HIPCHK(c, dalloc(&d.op, gm));
    HIPCHK(c, dalloc(&d.op_alt, gm));
    HIPCHK(c, dalloc(&d.vp, gm));
    // exit waves (K x g frames), then tmp (max(g, K) frames), then tmp2 (g frames): the position
    // step's transforms of tmp and tmp2 share the exit waves' FFT launches
    HIPCHK(c, dalloc(&d.psi, gm * d.k + (tcap + g) * mm));
    d.tmp = d.psi + gm * d.k;
    d.tmp2 = d.tmp + tcap * mm;
    HIPCHK(c, dalloc(&d.pu, gm));
    HIPCHK(c, dalloc(&d.gscal, (size_t)MLS_S_COEF + g));
    HIPCHK(c, dalloc(&d.ftot, (size_t)8 * g));
    HIPCHK(c, dalloc(&d.dxb, gm));
    HIPCHK(c, dalloc(&d.dxb_alt, gm));
    HIPCHK(c, dalloc(&d.pos_part, (size_t)4 * g * (d.m / 16)));
    HIPCHK(c, dalloc(&d.pos_ftot, (size_t)4 * g));
    HIPCHK(c, dalloc(&d.pum_part, (size_t)mls_pum_part_len(g, d.k, mm)));
    HIPCHK(c, dalloc(&d.vp_part, (size_t)g * ch + 2 * nb));
    HIPCHK(c, dalloc(&d.vp_ftot, (size_t)8 * g));
    HIPCHK(c, dalloc(&d.gsel, (size_t)1 + 2 * g));
    d.gmax = g;
  }
  const long need = std::max<long>(
      {6L * g * ch, (long)g * ch + 2 * nb, 2L * std::max(d.rows, d.cols), 1024L, mls_part_min(d.k)});
  if (need > d.part_cap) {
    mls_drop_graphs(d);
    if (d.part) hipFree(d.part);
    d.part = nullptr;
    d.part_cap = 0;
    HIPCHK(c, dalloc(&d.part, (size_t)need));
    d.part_cap = need;
  }
  return PTYK_OK;
}

int ptyk_mls_begin(ptyk_ctx* c, const ptyk_mls_config* cfg, const float* positions, const float* variable_mode,
                   const double* evolution_draws, const double* frame_power) {
  if (!c || !cfg || !positions || !variable_mode || !evolution_draws) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->have_frames || !c->have_probe) return fail(c, PTYK_ESTATE, "frames and probe must be set");
  hipSetDevice(c->cfg.device);
  const int n = c->cfg.n_frames, m = c->cfg.frame_size, rows = c->cfg.obj_rows, cols = c->cfg.obj_cols;
  if (cfg->n_objects != 1) return fail(c, PTYK_EINVAL, "n_objects must be 1");
  if (cfg->n_rec < 0 || cfg->delta_p < 0 || !(cfg->beta_lsq >= 0))
    return fail(c, PTYK_EINVAL, "bad MLs parameters");
  const int* b = cfg->object_box;
  if (b[0] < 1 || b[1] < b[0] || b[1] > rows || b[2] < 1 || b[3] < b[2] || b[3] > cols)
    return fail(c, PTYK_EINVAL, "object_box must lie inside the canvas (1-based, inclusive)");
  for (int i = 0; i < 2 * n; ++i)
    if (!std::isfinite(positions[i])) return fail(c, PTYK_EINVAL, "non-finite position");
  std::vector<double> fp(n);
  if (frame_power) {
    for (int i = 0; i < n; ++i) {
      if (!(frame_power[i] > 0) || !std::isfinite(frame_power[i]))
        return fail(c, PTYK_EINVAL, "frame_power must be finite and > 0");
      fp[i] = frame_power[i];
    }
  } else {
    std::vector<float> a((size_t)n * c->mm);
    int r = ptyk_get_frames_f32(c, a.data());
    if (r) return r;
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (long q = 0; q < c->mm; ++q) s += (double)a[(size_t)i * c->mm + q] * a[(size_t)i * c->mm + q];
      if (!(s > 0)) return fail(c, PTYK_EINVAL, "frame " + std::to_string(i) + " is all zero");
      fp[i] = s;
    }
  }
  free_mls(c);
  MlsDev& d = c->mls;
  d.n = n;
  d.m = m;
  d.k = c->cfg.n_modes;
  d.rows = rows;
  d.cols = cols;
  d.pps = cfg->probe_pos_search;
  d.probe_reconstruct = cfg->probe_reconstruct;
  d.object_reconstruct = cfg->object_reconstruct;
  d.variable_probe = cfg->variable_probe;
  d.variable_intensity = cfg->variable_intensity;
  d.variable_probe_modes = cfg->variable_probe_modes;
  d.apply_subpix_shift = cfg->apply_subpix_shift;
  d.apply_multimodal_update = cfg->apply_multimodal_update;
  d.n_objects = cfg->n_objects;
  d.n_rec = cfg->n_rec;
  for (int i = 0; i < 4; ++i) d.box[i] = b[i];
  d.delta_p = cfg->delta_p;
  d.beta_lsq = cfg->beta_lsq;
  d.beta_object = cfg->beta_object;
  d.beta_probe = cfg->beta_probe;
  d.obj = c->obj;
  d.probe = c->probe;
  d.amp = c->amp;
  d.amp_scale = c->amp_scale;
  d.tw = c->tw;
  const long canvas = (long)rows * cols;
  HIPCHK(c, dalloc(&d.vmode, c->mm));
  HIPCHK(c, dalloc(&d.evo, 2 * (size_t)n));
  HIPCHK(c, dalloc(&d.illum, canvas));
  HIPCHK(c, dalloc(&d.pos, n));
  HIPCHK(c, dalloc(&d.pos_o, n));
  HIPCHK(c, dalloc(&d.roi, n));
  HIPCHK(c, dalloc(&d.sub, n));
  HIPCHK(c, dalloc(&d.step_obj, n));
  HIPCHK(c, dalloc(&d.step_probe, n));
  HIPCHK(c, dalloc(&d.ferr, n));
  HIPCHK(c, dalloc(&d.flag, 1));
  HIPCHK(c, dalloc(&d.spec, 2 * c->mm));
  HIPCHK(c, dalloc(&d.pum, c->mm));
  HIPCHK(c, dalloc(&d.acc, c->mm));
  HIPCHK(c, dalloc(&d.precond, canvas));
  HIPCHK(c, dalloc(&d.gframes, n));
  HIPCHK(c, dalloc(&d.goffs_dev, (size_t)n + 1));
  int r = mls_reserve(c, 1);
  if (r) return r;
  std::vector<double> evo(2 * (size_t)n, 1.0);
  for (int i = 0; i < n; ++i) evo[n + i] = evolution_draws[i];
  HIPCHK(c, hipMemcpyAsync(d.pos, positions, n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.pos_o, positions, n * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.vmode, variable_mode, c->mm * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.evo, evo.data(), evo.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, mls_roi(d, c->stream));
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, d.flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flag) {
    free_mls(c);
    return fail(c, PTYK_ERANGE, "a scan position puts its ROI outside the object canvas");
  }
  HIPCHK(c, mls_init(d, fp.data(), c->stream));
  c->mls_ready = true;
  return PTYK_OK;
}

static int mls_run(ptyk_ctx* c, int32_t iter, const int32_t* group_frames, const int32_t* group_offsets,
                   int32_t n_groups, float* fourier_errors, float* kernel_ms) {
  if (!c || !group_frames || !group_offsets) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->mls_ready) return fail(c, PTYK_ESTATE, "ptyk_mls_begin first");
  if (iter < 1 || n_groups < 0) return fail(c, PTYK_EINVAL, "iter must be >= 1");
  hipSetDevice(c->cfg.device);
  MlsDev& d = c->mls;
  const int n = d.n;
  if (group_offsets[0] != 0) return fail(c, PTYK_EINVAL, "group_offsets[0] must be 0");
  int gmax = 0;
  for (int g = 0; g < n_groups; ++g) {
    const int len = group_offsets[g + 1] - group_offsets[g];
    if (len <= 0) return fail(c, PTYK_EINVAL, "empty group");
    gmax = std::max(gmax, len);
  }
  const int total = group_offsets[n_groups];
  if (total > n) return fail(c, PTYK_EINVAL, "more frames than positions in the groups");
  std::vector<char> seen(n, 0);
  for (int i = 0; i < total; ++i) {
    const int f = group_frames[i];
    if (f < 0 || f >= n || seen[f]) return fail(c, PTYK_EINVAL, "group frames must be distinct indices < n");
    seen[f] = 1;
  }
  int r = mls_reserve(c, std::max(gmax, 1));
  if (r) return r;
  if (total > 0)
    HIPCHK(c, hipMemcpyAsync(d.gframes, group_frames, total * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.goffs_dev, group_offsets, (n_groups + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
  d.mod_cb = c->mls_mod_cb;
  d.pum_fpc = c->mls_pum_fpc;
  d.row_waves = c->mls_row_waves != 0;
  d.use_graphs = c->mls_graphs != 0;
  MlsTimer tm;
  if (kernel_ms) {
    // marks, then start/stop pairs for the three streamed passes (at most four launches a group)
    const size_t marks = (size_t)24 * n_groups + 64, pairs = (size_t)4 * n_groups + 8;
    r = timer_reserve(c, marks + 2 * pairs);
    if (r) return r;
    tm.ev = c->evpool.data();
    tm.cat = c->evcat.data();
    tm.cap = (int)marks;
    tm.kev = c->evpool.data() + marks;
    tm.kcat = c->evcat.data() + marks;
    tm.kcap = (int)pairs;
  }
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(c, mls_iteration(d, iter, group_offsets, n_groups, c->stream, kernel_ms ? &tm : nullptr));
  const double host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, d.flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (fourier_errors && iter < 10)
    HIPCHK(c, hipMemcpyAsync(fourier_errors, d.ferr, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // also keeps group_frames alive until the copy ran
  if (kernel_ms) {
    for (int i = 0; i <= MLS_T_COUNT; ++i) kernel_ms[i] = 0.f;
    for (int i = 0; i + 1 < tm.n; ++i) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, tm.ev[i], tm.ev[i + 1]));
      if (tm.cat[i] >= 0 && tm.cat[i] < MLS_T_COUNT) kernel_ms[tm.cat[i]] += ms;
    }
    for (int i = 0; i < tm.kn; ++i) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, tm.kev[2 * i], tm.kev[2 * i + 1]));
      if (tm.kcat[i] >= 0 && tm.kcat[i] < MLS_T_COUNT) kernel_ms[tm.kcat[i]] += ms;
    }
    kernel_ms[MLS_T_COUNT] = (float)host_ms;
  }
  if (flag) {
    HIPCHK(c, hipMemsetAsync(d.flag, 0, sizeof(int), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return fail(c, PTYK_ERANGE, "refined positions moved a ROI outside the object canvas (clamped)");
  }
  return PTYK_OK;
}

int ptyk_mls_iteration(ptyk_ctx* c, int32_t iter, const int32_t* group_frames, const int32_t* group_offsets,
                       int32_t n_groups, float* fourier_errors) {
  return mls_run(c, iter, group_frames, group_offsets, n_groups, fourier_errors, nullptr);
}

int ptyk_mls_iteration_timed(ptyk_ctx* c, int32_t iter, const int32_t* group_frames, const int32_t* group_offsets,
                             int32_t n_groups, float* fourier_errors, float* kernel_ms) {
  if (!kernel_ms) return fail(c, PTYK_EINVAL, "null argument");
  return mls_run(c, iter, group_frames, group_offsets, n_groups, fourier_errors, kernel_ms);
}

static int mls_state_copy(ptyk_ctx* c, const ptyk_mls_state* st, bool out) {
  if (!c || !st) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->mls_ready) return fail(c, PTYK_ESTATE, "ptyk_mls_begin first");
  hipSetDevice(c->cfg.device);
  MlsDev& d = c->mls;
  const size_t n = d.n;
  const hipMemcpyKind kind = out ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
  auto cp = [&](void* host, void* dev, size_t bytes) -> hipError_t {
    if (!host) return hipSuccess;
    return out ? hipMemcpyAsync(host, dev, bytes, kind, c->stream) : hipMemcpyAsync(dev, host, bytes, kind, c->stream);
  };
  HIPCHK(c, cp(st->positions, d.pos, n * sizeof(float2)));
  HIPCHK(c, cp(st->positions_start, d.pos_o, n * sizeof(float2)));
  HIPCHK(c, cp(st->roi, d.roi, n * sizeof(int2)));
  HIPCHK(c, cp(st->subpixel, d.sub, n * sizeof(float2)));
  HIPCHK(c, cp(st->evolution, d.evo, 2 * n * sizeof(double)));
  HIPCHK(c, cp(st->variable_mode, d.vmode, c->mm * sizeof(float2)));
  HIPCHK(c, cp(st->illumination, d.illum, (size_t)d.rows * d.cols * sizeof(float)));
  HIPCHK(c, cp(st->beta_object, d.step_obj, n * sizeof(float2)));
  HIPCHK(c, cp(st->beta_probe, d.step_probe, n * sizeof(float2)));
  HIPCHK(c, cp(st->fourier_errors, d.ferr, n * sizeof(float)));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

int ptyk_mls_get_state(ptyk_ctx* c, ptyk_mls_state* st) {
  int r = mls_state_copy(c, st, true);
  if (r == PTYK_OK) st->max_illumination = c->mls.max_illum;
  return r;
}

int ptyk_mls_set_state(ptyk_ctx* c, const ptyk_mls_state* st) {
  if (st && st->roi) {
    const MlsDev& d = c->mls;
    for (int i = 0; i < d.n; ++i) {
      const int ox = st->roi[2 * i], oy = st->roi[2 * i + 1];
      if (ox < 0 || oy < 0 || ox + d.m > d.rows || oy + d.m > d.cols)
        return fail(c, PTYK_ERANGE, "ROI outside the object canvas");
    }
  }
  int r = mls_state_copy(c, st, false);
  if (r == PTYK_OK && st->max_illumination > 0) c->mls.max_illum = (float)st->max_illumination;
  return r;
}

// ---------------------------------------------------------------- batched scans
// Several reconstructions advanced together: one DM iteration of every scan is captured once per
// flag set into a HIP graph with one branch per scan (fork/join through events on a capture
// stream) and replayed on later calls, so an iteration of the whole batch is one graph launch.
// Each branch is exactly that scan's own launch sequence, so every scan's result is bitwise the
// one of its standalone run.
struct ptyk_batch {
  std::vector<ptyk_ctx*> ctx;
  int device = 0;
  hipStream_t s0 = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  struct Graph {
    uint32_t flags;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  std::vector<uintptr_t> sig;  // buffers and launch shapes the graphs were captured with
  std::string err;
};

static std::vector<uintptr_t> dm_signature(const ptyk_ctx* c) {
  return {(uintptr_t)c->psi, (uintptr_t)c->row_ptr, (uintptr_t)c->row_frames, (uintptr_t)c->probe_partial,
          (uintptr_t)c->obj_partial, (uintptr_t)c->probe_partial_sum, (uintptr_t)c->d_obj, (uintptr_t)c->d_probe,
          (uintptr_t)c->wave_row_ptr, (uintptr_t)c->wave_row_frames, (uintptr_t)c->dm_wave, (uintptr_t)c->dm_waves,
          (uintptr_t)c->probe_chunks, (uintptr_t)c->frames_per_chunk, (uintptr_t)dm_wave_frames(c),
          (uintptr_t)c->dm_gen};
}

static void batch_drop_graphs(ptyk_batch* b) {
  for (auto& g : b->graphs) hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

#define BCHK(b, call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
      return PTYK_EDEVICE;                                                   \
    }                                                                        \
  } while (0)

int ptyk_batch_create(ptyk_ctx* const* ctxs, int32_t n, ptyk_batch** out) {
  if (!ctxs || !out || n <= 0) return fail(nullptr, PTYK_EINVAL, "bad batch arguments");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(nullptr, PTYK_EINVAL, "null context in batch");
    if (ctxs[i]->cfg.device != ctxs[0]->cfg.device) return fail(nullptr, PTYK_EINVAL, "contexts on different devices");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(nullptr, PTYK_EINVAL, "a context appears twice in the batch");
  }
  ptyk_batch* b = new ptyk_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->device = ctxs[0]->cfg.device;
  hipSetDevice(b->device);
  hipError_t e = hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  b->join.assign(n, nullptr);
  for (int i = 0; i < n && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    ptyk_batch_destroy(b);
    return fail(nullptr, PTYK_EDEVICE, std::string("batch setup: ") + hipGetErrorString(e));
  }
  *out = b;
  return PTYK_OK;
}

void ptyk_batch_destroy(ptyk_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->s0) hipStreamSynchronize(b->s0);
  batch_drop_graphs(b);
  for (auto e : b->join)
    if (e) hipEventDestroy(e);
  if (b->fork) hipEventDestroy(b->fork);
  if (b->done) hipEventDestroy(b->done);
  if (b->s0) hipStreamDestroy(b->s0);
  delete b;
}

const char* ptyk_batch_last_error(const ptyk_batch* b) { return b ? b->err.c_str() : g_create_err.c_str(); }

int ptyk_batch_dm_iteration(ptyk_batch* b, uint32_t flags) {
  if (!b) return PTYK_EINVAL;
  hipSetDevice(b->device);
  // host-side preparation of every scan (allocations, frame lists, first-call factors) runs
  // eagerly; a scan whose DM state is not started yet runs this iteration eagerly too
  bool eager = false;
  for (ptyk_ctx* c : b->ctx) {
    if (!c->have_frames || !c->have_pos || !c->have_probe) {
      b->err = "frames, positions and probe must be set in every scan";
      return PTYK_ESTATE;
    }
    if (!c->dm_started || !c->psi) eager = true;
    int r = dm_prepare(c);
    if (r) {
      b->err = c->err;
      return r;
    }
  }
  if (eager) {
    for (ptyk_ctx* c : b->ctx) {
      int r = dm_launch(c, flags, nullptr);
      if (r) {
        b->err = c->err;
        return r;
      }
    }
    return PTYK_OK;
  }
  std::vector<uintptr_t> sig;
  for (ptyk_ctx* c : b->ctx) {
    auto v = dm_signature(c);
    sig.insert(sig.end(), v.begin(), v.end());
  }
  if (sig != b->sig) {
    batch_drop_graphs(b);
    b->sig = sig;
  }
  hipGraphExec_t exec = nullptr;
  for (auto& g : b->graphs)
    if (g.flags == flags) exec = g.exec;
  if (!exec) {
    // the scans' own streams must be idle before their work moves into the graph
    for (ptyk_ctx* c : b->ctx) BCHK(b, hipStreamSynchronize(c->stream));
    BCHK(b, hipStreamBeginCapture(b->s0, hipStreamCaptureModeRelaxed));
    BCHK(b, hipEventRecord(b->fork, b->s0));
    int rc = PTYK_OK;
    for (size_t i = 0; i < b->ctx.size() && rc == PTYK_OK; ++i) {
      ptyk_ctx* c = b->ctx[i];
      if (hipStreamWaitEvent(c->stream, b->fork, 0) != hipSuccess) rc = PTYK_EDEVICE;
      if (rc == PTYK_OK) rc = dm_launch(c, flags, nullptr);
      if (rc == PTYK_OK && hipEventRecord(b->join[i], c->stream) != hipSuccess) rc = PTYK_EDEVICE;
      if (rc == PTYK_OK && hipStreamWaitEvent(b->s0, b->join[i], 0) != hipSuccess) rc = PTYK_EDEVICE;
    }
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(b->s0, &graph);
    if (rc != PTYK_OK || ec != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      b->err = "graph capture failed";
      return rc != PTYK_OK ? rc : PTYK_EDEVICE;
    }
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      b->err = std::string("graph instantiate: ") + hipGetErrorString(ei);
      return PTYK_EDEVICE;
    }
    b->graphs.push_back({flags, exec});
  }
  // the replay follows whatever is already queued on each scan's own stream (a standalone
  // iteration, a copy): the capture path synchronised them, a replay must order itself
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    BCHK(b, hipEventRecord(b->join[i], b->ctx[i]->stream));
    BCHK(b, hipStreamWaitEvent(b->s0, b->join[i], 0));
  }
  BCHK(b, hipGraphLaunch(exec, b->s0));
  // later calls on any scan's stream see this iteration
  BCHK(b, hipEventRecord(b->done, b->s0));
  for (ptyk_ctx* c : b->ctx) BCHK(b, hipStreamWaitEvent(c->stream, b->done, 0));
  return PTYK_OK;
}

int ptyk_batch_synchronize(ptyk_batch* b) {
  if (!b) return PTYK_EINVAL;
  hipSetDevice(b->device);
  BCHK(b, hipStreamSynchronize(b->s0));
  for (ptyk_ctx* c : b->ctx) BCHK(b, hipStreamSynchronize(c->stream));
  return PTYK_OK;
}

int ptyk_batch_graph_count(const ptyk_batch* b) { return b ? (int)b->graphs.size() : -1; }

// ---------------------------------------------------------------- object error
int ptyk_set_true_object(ptyk_ctx* c, const float* truth) {
  if (!c || !truth) return fail(c, PTYK_EINVAL, "null argument");
  hipSetDevice(c->cfg.device);
  const long canvas = (long)c->cfg.obj_rows * c->cfg.obj_cols;
  if (!c->truth) {
    HIPCHK(c, dalloc(&c->truth, canvas));
    HIPCHK(c, dalloc(&c->cov, canvas));
    if (!c->pint) HIPCHK(c, dalloc(&c->pint, c->mm));
    HIPCHK(c, dalloc(&c->epart, 5 * 1024));
  }
  return copy_in(c, c->truth, truth, canvas * sizeof(float2));
}

int ptyk_object_error(ptyk_ctx* c, double frac, double* out) {
  if (!c || !out) return fail(c, PTYK_EINVAL, "null argument");
  if (!c->truth) return fail(c, PTYK_ESTATE, "ptyk_set_true_object first");
  if (!(frac >= 0 && frac < 1)) return fail(c, PTYK_EINVAL, "frac must be in [0, 1)");
  const int2* roi = c->mls_ready ? c->mls.roi : (c->have_pos ? c->roi_dev : nullptr);
  if (!roi) return fail(c, PTYK_ESTATE, "no scan positions");
  hipSetDevice(c->cfg.device);
  HIPCHK(c, object_error(c->obj, c->truth, c->probe, c->cfg.n_modes, c->cfg.frame_size, roi, c->cfg.n_frames,
                         c->cfg.obj_rows, c->cfg.obj_cols, frac, c->cov, c->pint, c->epart, 5 * 1024, out, c->stream));
  return PTYK_OK;
}

// Learned warm start: the object becomes the probe-intensity-weighted average of the given
// (amplitude, phase) patches placed at the ROIs of the selected frames (MLs ROIs after
// ptyk_mls_begin, otherwise the positions' ROIs).
int ptyk_stitch_patches(ptyk_ctx* c, const float* patches, const int32_t* frames, int64_t n, float rel_threshold,
                        float fill_re, float fill_im) {
  if (!c || !patches || !frames || n <= 0 || n > c->cfg.n_frames || !(rel_threshold >= 0.f))
    return fail(c, PTYK_EINVAL, "bad argument");
  const int2* roi = c->mls_ready ? c->mls.roi : (c->have_pos ? c->roi_dev : nullptr);
  if (!roi) return fail(c, PTYK_ESTATE, "no scan positions");
  for (int64_t i = 0; i < n; ++i)
    if (frames[i] < 0 || frames[i] >= c->cfg.n_frames) return fail(c, PTYK_EINVAL, "frame index out of range");
  hipSetDevice(c->cfg.device);
  if (!c->pint) HIPCHK(c, dalloc(&c->pint, c->mm));
  float* d_p = nullptr;
  int* d_sel = nullptr;
  const size_t pbytes = (size_t)n * 2 * c->mm * sizeof(float);
  hipError_t e = hipMalloc((void**)&d_p, pbytes);
  if (e == hipSuccess) e = hipMalloc((void**)&d_sel, n * sizeof(int));
  if (e == hipSuccess) e = hipMemcpyAsync(d_p, patches, pbytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_sel, frames, n * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = stitch_patches(c->obj, d_p, c->probe, c->cfg.n_modes, c->cfg.frame_size, roi, d_sel, (int)n, c->cfg.obj_rows,
                       c->cfg.obj_cols, c->pint, rel_threshold, make_float2(fill_re, fill_im), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_p) hipFree(d_p);
  if (d_sel) hipFree(d_sel);
  HIPCHK(c, e);
  return PTYK_OK;
}

// ---------------------------------------------------------------- utilities
int ptyk_fft2(ptyk_ctx* c, float* data, int64_t n, int inverse) {
  if (!c || !data || n < 0) return fail(c, PTYK_EINVAL, "bad argument");
  hipSetDevice(c->cfg.device);
  const size_t bytes = (size_t)n * c->mm * sizeof(float2);
  float2* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, bytes ? bytes : 8));
  hipError_t e = hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = fft2_batch(d, n, c->cfg.frame_size, inverse != 0, c->tw, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  HIPCHK(c, e);
  return PTYK_OK;
}

int ptyk_preprocess_frames(ptyk_ctx* c, const float* raw, int64_t count, const ptyk_ingest_params* p, float* out) {
  if (!c) return PTYK_EINVAL;
  if (!raw || !p || !out || count < 0 || p->raw_rows <= 0 || p->raw_cols <= 0) return fail(c, PTYK_EINVAL, "bad argument");
  if (hipSetDevice(c->cfg.device) != hipSuccess) return fail(c, PTYK_EDEVICE, "hipSetDevice failed");
  if (count == 0) return PTYK_OK;
  IngestArgs a{};
  a.raw_rows = p->raw_rows;
  a.raw_cols = p->raw_cols;
  a.size = c->cfg.frame_size;
  a.center_row = p->center_row;
  a.center_col = p->center_col;
  a.rotate = (p->rotate90 % 4 == 1) ? 1 : 0;
  a.fft_shift = p->fft_shift ? 1 : 0;
  a.threshold_sqrt = p->threshold_sqrt ? 1 : 0;
  a.square_root = p->square_root ? 1 : 0;
  a.threshold = p->threshold;
  const size_t nraw = (size_t)count * p->raw_rows * p->raw_cols, nout = (size_t)count * c->mm;
  float *d_raw = nullptr, *d_out = nullptr;
  hipError_t e = hipMalloc((void**)&d_raw, nraw * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&d_out, nout * sizeof(float));
  if (e == hipSuccess) e = hipMemcpyAsync(d_raw, raw, nraw * sizeof(float), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = ingest_frames(d_raw, d_out, count, a, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, nout * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_raw) hipFree(d_raw);
  if (d_out) hipFree(d_out);
  HIPCHK(c, e);
  return PTYK_OK;
}

int ptyk_simulate_noalloc(ptyk_ctx* c, const float* obj, int32_t rows, int32_t cols, const float* probe,
                          const float* pos, int64_t n, int32_t squared, float* out) {
  if (!c) return PTYK_EINVAL;
  if (!obj || !probe || !pos || !out || rows <= 0 || cols <= 0 || n < 0) return fail(c, PTYK_EINVAL, "bad argument");
  const int m = c->cfg.frame_size;
  const size_t need = (size_t)n * c->mm * sizeof(float2);
  if (need > noalloc_scratch_bytes())
    return fail(c, PTYK_EINVAL,
                std::to_string(n) + " exit waves of " + std::to_string(m) + "^2 need " + std::to_string(need) +
                    " bytes of scratch; the fixed array holds " + std::to_string(noalloc_scratch_bytes()));
  if (hipSetDevice(c->cfg.device) != hipSuccess) return fail(c, PTYK_EDEVICE, "hipSetDevice failed");
  float2* psi = static_cast<float2*>(noalloc_scratch(c->cfg.device));
  if (!psi) return fail(c, PTYK_ESTATE, "scratch array not set up on this device");
  if (n == 0) return PTYK_OK;
  HIPCHK(c, simulate_exit_waves(reinterpret_cast<const float2*>(obj), rows, cols,
                                reinterpret_cast<const float2*>(probe), reinterpret_cast<const float2*>(pos), psi, n,
                                m, c->cfg.tex_quantize, c->stream));
  HIPCHK(c, fft2_batch(psi, n, m, false, c->tw, c->stream));
  HIPCHK(c, abs_frames(psi, out, n * c->mm, squared, c->stream));
  return PTYK_OK;
}

int ptyk_simulate(ptyk_ctx* c, const float* obj, int32_t rows, int32_t cols, const float* probe, const float* pos,
                  int64_t n, int32_t squared, float* out) {
  if (!c || !obj || !probe || !pos || !out || rows <= 0 || cols <= 0 || n < 0)
    return fail(c, PTYK_EINVAL, "bad argument");
  hipSetDevice(c->cfg.device);
  const int m = c->cfg.frame_size;
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(pos[2 * i]) || !std::isfinite(pos[2 * i + 1]))
      return fail(c, PTYK_EINVAL, "non-finite position");
  const size_t canvas = (size_t)rows * cols;
  float2 *d_obj = nullptr, *d_probe = nullptr, *d_pos = nullptr, *d_psi = nullptr;
  float* d_out = nullptr;
  hipError_t e = hipMalloc((void**)&d_obj, canvas * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc((void**)&d_probe, c->mm * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc((void**)&d_pos, (n ? n : 1) * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc((void**)&d_psi, (n ? n : 1) * c->mm * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc((void**)&d_out, (n ? n : 1) * c->mm * sizeof(float));
  if (e == hipSuccess) e = hipMemcpyAsync(d_obj, obj, canvas * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_probe, probe, c->mm * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(d_pos, pos, n * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && n)
    e = simulate_exit_waves(d_obj, rows, cols, d_probe, d_pos, d_psi, n, m, c->cfg.tex_quantize, c->stream);
  if (e == hipSuccess && n) e = fft2_batch(d_psi, n, m, false, c->tw, c->stream);
  if (e == hipSuccess && n) e = abs_frames(d_psi, d_out, n * c->mm, squared, c->stream);
  if (e == hipSuccess && n) e = hipMemcpyAsync(out, d_out, n * c->mm * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (void* p : {(void*)d_obj, (void*)d_probe, (void*)d_pos, (void*)d_psi, (void*)d_out})
    if (p) hipFree(p);
  HIPCHK(c, e);
  return PTYK_OK;
}

}  // extern "C"
