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

#ifdef __cplusplus
}
#endif
#endif /* PTYCHO_H */
