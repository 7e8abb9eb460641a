/*
 * ptycho.h -- C ABI of the host-side driver (scan geometry, configuration, the
 * IPhaser / IPhasingMethod layer).  The GPU engine underneath is declared in ptyk.h.
 */
#ifndef PTYCHO_H
#define PTYCHO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sample-plane pixel size lambda*z/(size*dx_d), rounded to float as the reference's
 * Parameters::samplePlanePixelSize returns real_t (Parameters.h:175-179). */
double ptycho_sample_pixel_size(double lambda, double z, int32_t size, double dx_d);

typedef struct {
  int32_t nx, ny;          /* scan grid (scanDims.x rows, scanDims.y columns) */
  double step_x, step_y;   /* metres; stored as float like ExperimentParams::stepSize */
  double dx_s;             /* sample-plane pixel size (metres) */
  double drift_x, drift_y; /* metres per step (ExperimentParams::drift) */
  int32_t jitter_radius;   /* pixels; 0 = none */
  int32_t mirror_x, mirror_y, flip;
  int32_t probe_size;      /* M */
  int32_t reserved;
  uint64_t seed;           /* jitter stream */
} ptycho_mesh_params;

/* Cartesian scan (CartesianScanMesh::generateMesh, CartesianScanMesh.cpp:57-92; d_generateMesh<1>,
 * ScanMeshKernels.cu:46-93): pos_out nx*ny x (row, col) float32 already shifted by the minima,
 * canvas_out[2] = ceil(max - min + M), extent_out[4] = {min_x, min_y, max_x, max_y} (nullable). */
int ptycho_mesh_cartesian(const ptycho_mesh_params* p, float* pos_out, int32_t* canvas_out, float* extent_out);

/* Position list given in metres (ListScanMesh, ListScanMesh.cpp:55-103): n x (x, y). */
int ptycho_mesh_list(const ptycho_mesh_params* p, const double* metres, int64_t n, float* pos_out,
                     int32_t* canvas_out, float* extent_out);
/* Spiral of nx*ny points (SpiralScanMesh::generateMesh, SpiralScanMesh.cpp:53-78). */
int ptycho_mesh_spiral(const ptycho_mesh_params* p, float* pos_out, int32_t* canvas_out, float* extent_out);
/* Cartesian positions before any shift, nx*ny x (row, col) pixels (d_generateMesh<1>). */
int ptycho_mesh_cartesian_raw(const ptycho_mesh_params* p, float* raw_out);

/* MLs object canvas and normalisation box for unshifted positions (CartesianScanMesh::generateMeshML,
 * CartesianScanMesh.cpp:94-177; IPtychoScanMesh::precalculate_ROI, IPtychoScanMesh.cpp:179-184):
 * extent = {min_x, min_y, max_x, max_y} over the scan buffer (padding included), canvas_out[2] =
 * ceil(1.2 x span of the rounded ROI bounds), box_out[4] = 1-based inclusive {row0, row1, col0, col1}. */
int ptycho_mls_canvas(const float* pos, int64_t n, const float* extent, int32_t probe_size, int32_t* canvas_out,
                      int32_t* box_out);
/* MLs frame groups (IPtychoScanMesh::get_nonoverlapping_indices, IPtychoScanMesh.cpp:826-913):
 * min(niter, 30) seeded random partitions of n frames into groups of max(1, n/100), largest first.
 * frames_out: partitions x n, offsets_out: partitions x (groups + 1); pass NULL to get the counts. */
int ptycho_mls_partitions(int64_t n, int32_t niter, uint64_t seed, int32_t* frames_out, int32_t* offsets_out,
                          int32_t* n_groups, int32_t* n_partitions);

/* Initial MLs probe modes (Probe::initMLH, Probe.cpp:376-494) with its random draws given:
 * out (K x M x M complex64) holds mode 0 on entry when it is non-zero (a guess), else all zeros
 * and mode 0 is the Fresnel-zone-plate model; modes 1..K-1 are mode 0 tilted by
 * exp(-2 pi i (draws[2i-2] (col/M - 0.5) + draws[2i-1] (row/M - 0.5))); all modes are then replaced
 * by the eigenmodes of their overlap matrix, strongest first.  engine: a ptyk_ctx of frame size M
 * (its FFT forms the zone-plate probe). */
int ptycho_mls_probe_modes(void* engine, int32_t m, int32_t k, double lambda, double dx_s, double beam_size,
                           const float* draws, float* out);

/* Compact MLs frame groups (IPtychoScanMesh::get_close_indices, IPtychoScanMesh.cpp:422-824, the
 * "MLC" grouping): k-means clusters of the positions (n x 2) into ceil(n / max(1, n/100)) groups,
 * balanced to at most two sizes and refined by size-preserving swaps toward the nearest median
 * centre; groups by decreasing size, frames ascending.  frames_out: n, offsets_out: groups + 1;
 * pass NULL to get the count. */
int ptycho_mls_compact_groups(const float* pos, int64_t n, uint64_t seed, int32_t* frames_out, int32_t* offsets_out,
                              int32_t* n_groups);

/* Complex CSV ("a+bj", ", "-separated rows) or raw binary, as Cuda2DArray::save / load
 * (Cuda2DArray.cpp:46-108, Cuda2DArray.hpp:439-464). */
int ptycho_save_complex_csv(const char* path, const float* c64, int32_t rows, int32_t cols, int32_t binary);
int ptycho_load_complex_csv(const char* path, float* out, int32_t max_elems, int32_t* rows, int32_t* cols);

/* ---------------------------------------------------------------- IPhaser
 * One reconstruction driver (IPhaser, IPhaser.h:135-184) with its own parameter set
 * (CXParams, Parameters.h:168-203).  All functions return 0 on success, -1 on failure
 * (message via ptycho_phaser_last_error); nothing exits the process. */
typedef struct ptycho_phaser ptycho_phaser;

int ptycho_phaser_create(ptycho_phaser** out);
void ptycho_phaser_destroy(ptycho_phaser* h);
const char* ptycho_phaser_last_error(const ptycho_phaser* h);

/* Parameters::parseFromCPython, one keyword at a time (Parameters.cpp:380-653; keyword names
 * of ptychopy.c:242-246 plus "algorithm", "RMS" and "seed"); value as text. */
int ptycho_phaser_set_param(ptycho_phaser* h, const char* key, const char* value);
/* Parameters::parseFromCommandLine on a "-key=value ..." string (Parameters.cpp:66-378). */
int ptycho_phaser_parse_cmdline(ptycho_phaser* h, const char* cmdline);
int ptycho_phaser_get_param(ptycho_phaser* h, const char* key, char* buf, int32_t len);

/* numpy inputs of the *np entry points (ptychopy.c:346-441): frames n x rows x cols float32,
 * positions n x 2 metres, object / probe guesses complex64. */
int ptycho_phaser_set_diffraction(ptycho_phaser* h, const float* data, int64_t n, int32_t rows, int32_t cols);
int ptycho_phaser_set_positions(ptycho_phaser* h, const double* xy, int64_t n);
int ptycho_phaser_set_object_guess(ptycho_phaser* h, const float* c64, int32_t rows, int32_t cols);
int ptycho_phaser_set_probe_guess(ptycho_phaser* h, const float* c64, int32_t rows, int32_t cols);

int ptycho_phaser_init(ptycho_phaser* h);                                       /* IPhaser::init */
int ptycho_phaser_add_method(ptycho_phaser* h, const char* name, uint32_t iterations); /* "ePIE" | "DM" | "MLs" */
int ptycho_phaser_phase(ptycho_phaser* h);                                      /* IPhaser::phase */
int ptycho_phaser_phaseinit(ptycho_phaser* h);                                  /* IPhaser::phaseinit */
int ptycho_phaser_phasestep(ptycho_phaser* h, uint32_t i);                      /* IPhaser::phasestepvis */
int ptycho_phaser_phasepost(ptycho_phaser* h);                                  /* IPhaser::phasepost */
int ptycho_phaser_write_results(ptycho_phaser* h, int32_t r);                   /* writeResultsToDisk */

int ptycho_phaser_object_shape(ptycho_phaser* h, int32_t* rows, int32_t* cols);
int ptycho_phaser_get_object(ptycho_phaser* h, float* out);                     /* complex64 rows x cols */
int ptycho_phaser_probe_shape(ptycho_phaser* h, int32_t* modes, int32_t* size);
int ptycho_phaser_get_probe(ptycho_phaser* h, float* out);                      /* complex64 modes*size x size */
int64_t ptycho_phaser_num_positions(ptycho_phaser* h);
int ptycho_phaser_get_positions(ptycho_phaser* h, float* out, int64_t n);        /* pixels, canvas frame */
int ptycho_phaser_get_errors(ptycho_phaser* h, double* out, int32_t n);          /* IPhaser::getPhaseErrors */
/* MLs per-frame Fourier errors of iteration iter (1-based; IPhaser::m_fourierErrors, filled for
 * iter < 10, MLS.cpp:234-237); n frames. */
int ptycho_phaser_get_fourier_errors(ptycho_phaser* h, uint32_t iter, float* out, int64_t n);
double ptycho_phaser_phase_seconds(ptycho_phaser* h);
/* The engine handle (ptyk.h) the driver runs on; valid after ptycho_phaser_init. */
void* ptycho_phaser_engine(ptycho_phaser* h);

#ifdef __cplusplus
}
#endif
#endif /* PTYCHO_H */
