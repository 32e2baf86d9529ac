/* libatmo_h5: Trixi-compatible HDF5 solution and restart files.
 *
 * Host-only companion of libatmo_hip (it links the HDF5 C library, which the
 * compute library does not need).  The file layout is the one Trixi and
 * TrixiAtmo write, so their readers (Trixi2Vtk, Trixi.load_restart_file) take
 * these files and the files they write can be read here:
 *   root attributes  ndims, equations, polydeg, n_vars, n_elements,
 *                    mesh_type, mesh_file, time, dt, timestep
 *   datasets         variables_1 .. variables_<n_vars>: one 1D FP64 vector
 *                    per variable over the nodes in Trixi's order, each with
 *                    a string attribute `name`
 * (src/callbacks_step/save_solution_2d_manifold_in_3d_cartesian.jl:44-120 for
 * P4estMesh{2} with covariant or 3D equations; Trixi's save_solution_file
 * and save_restart_file for P4estMesh{3} and StructuredMesh [ext], which
 * write the same attributes and datasets).  Integers are stored as 64-bit
 * little-endian, reals as IEEE FP64, strings as scalar fixed-length UTF-8
 * (the types HDF5.jl gives Int64, Float64 and String). */
#ifndef ATMO_H5_H
#define ATMO_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t ndims;
  int64_t polydeg;
  int64_t n_vars;
  int64_t n_elements;
  int64_t timestep;
  double time;
  double dt;
  char equations[128];   /* e.g. "SplitCovariantShallowWaterEquations2D" */
  char mesh_type[64];    /* e.g. "P4estMesh" */
  char mesh_file[256];   /* file name of the mesh file, may be empty */
} atmo_h5_attributes;

/* Write a solution (or restart) file, clobbering an existing one: the
 * attributes above and data[v * n_values + q] as variables_<v+1> with
 * attribute name = varnames[v] (attributes->n_vars variables).  0 on
 * success, negative on failure (message from atmo_h5_last_error). */
int atmo_h5_write_solution(const char* filename, const atmo_h5_attributes* attributes,
                           int64_t n_values, const double* data, const char* const* varnames);

/* Read the root attributes of a solution or restart file. */
int atmo_h5_read_attributes(const char* filename, atmo_h5_attributes* attributes);

/* Read variables_<v> (1-based) into data[n_values]; its `name` attribute into
 * name (NUL-terminated, truncated to name_capacity, may be NULL).  The
 * dataset must hold exactly n_values values. */
int atmo_h5_read_variable(const char* filename, int32_t v, int64_t n_values, double* data,
                          char* name, int32_t name_capacity);

const char* atmo_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
