// Trixi-compatible HDF5 solution and restart files over the HDF5 C library:
// the root attributes and variables_<v> datasets that Trixi's and
// TrixiAtmo's save_solution_file write
// (src/callbacks_step/save_solution_2d_manifold_in_3d_cartesian.jl:74-97),
// with the HDF5 types HDF5.jl uses for Int64, Float64 and String.
#include <hdf5.h>

#include <cstring>
#include <string>

#include "../../../include/atmo_h5.h"

namespace {

thread_local std::string g_error;

int h5_fail(const std::string& msg) {
  g_error = msg;
  return -1;
}

// closes an HDF5 identifier of the given kind when it goes out of scope
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
};

bool write_attr_int(hid_t loc, const char* name, int64_t value) {
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id attr(H5Acreate2(loc, name, H5T_STD_I64LE, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return space.ok() && attr.ok() && H5Awrite(attr.id, H5T_NATIVE_INT64, &value) >= 0;
}

bool write_attr_double(hid_t loc, const char* name, double value) {
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id attr(H5Acreate2(loc, name, H5T_IEEE_F64LE, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return space.ok() && attr.ok() && H5Awrite(attr.id, H5T_NATIVE_DOUBLE, &value) >= 0;
}

// fixed-length UTF-8 string of the string's byte length (at least 1), NUL
// terminated when shorter, in a scalar dataspace
bool write_attr_string(hid_t loc, const char* name, const char* value) {
  const size_t len = std::strlen(value);
  Id type(H5Tcopy(H5T_C_S1), H5Tclose);
  if (!type.ok() || H5Tset_size(type.id, len > 0 ? len : 1) < 0 ||
      H5Tset_cset(type.id, H5T_CSET_UTF8) < 0)
    return false;
  Id space(H5Screate(H5S_SCALAR), H5Sclose);
  Id attr(H5Acreate2(loc, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  std::string buf(value);
  if (buf.empty()) buf.push_back('\0');
  return space.ok() && attr.ok() && H5Awrite(attr.id, type.id, buf.data()) >= 0;
}

bool read_attr_int(hid_t loc, const char* name, int64_t* out) {
  Id attr(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  return attr.ok() && H5Aread(attr.id, H5T_NATIVE_INT64, out) >= 0;
}

bool read_attr_double(hid_t loc, const char* name, double* out) {
  Id attr(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  return attr.ok() && H5Aread(attr.id, H5T_NATIVE_DOUBLE, out) >= 0;
}

// a fixed- or variable-length string attribute into out[cap]
bool read_attr_string(hid_t loc, const char* name, char* out, size_t cap) {
  Id attr(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  if (!attr.ok() || cap == 0) return false;
  Id ftype(H5Aget_type(attr.id), H5Tclose);
  if (!ftype.ok() || H5Tget_class(ftype.id) != H5T_STRING) return false;
  // the memory type keeps the file's character set (HDF5 converts between
  // string sizes, not between character sets)
  const H5T_cset_t cset = H5Tget_cset(ftype.id);
  if (H5Tis_variable_str(ftype.id) > 0) {
    Id mtype(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mtype.id, H5T_VARIABLE);
    H5Tset_cset(mtype.id, cset);
    char* s = nullptr;
    if (H5Aread(attr.id, mtype.id, &s) < 0) return false;
    std::strncpy(out, s ? s : "", cap - 1);
    out[cap - 1] = '\0';
    H5free_memory(s);
    return true;
  }
  const size_t len = H5Tget_size(ftype.id);
  std::string buf(len + 1, '\0');
  Id mtype(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(mtype.id, len + 1);
  H5Tset_strpad(mtype.id, H5T_STR_NULLTERM);
  H5Tset_cset(mtype.id, cset);
  if (H5Aread(attr.id, mtype.id, &buf[0]) < 0) return false;
  std::strncpy(out, buf.c_str(), cap - 1);
  out[cap - 1] = '\0';
  return true;
}

}  // namespace

extern "C" {

int atmo_h5_write_solution(const char* filename, const atmo_h5_attributes* a, int64_t n_values,
                           const double* data, const char* const* varnames) {
  if (!filename || !a || n_values < 0 || a->n_vars < 0 || (a->n_vars > 0 && n_values > 0 && !data) ||
      (a->n_vars > 0 && !varnames))
    return h5_fail("bad arguments");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  Id file(H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return h5_fail(std::string("cannot create ") + filename);
  const bool ok =
      write_attr_int(file.id, "ndims", a->ndims) &&
      write_attr_string(file.id, "equations", a->equations) &&
      write_attr_int(file.id, "polydeg", a->polydeg) && write_attr_int(file.id, "n_vars", a->n_vars) &&
      write_attr_int(file.id, "n_elements", a->n_elements) &&
      write_attr_string(file.id, "mesh_type", a->mesh_type) &&
      write_attr_string(file.id, "mesh_file", a->mesh_file) &&
      write_attr_double(file.id, "time", a->time) && write_attr_double(file.id, "dt", a->dt) &&
      write_attr_int(file.id, "timestep", a->timestep);
  if (!ok) return h5_fail("writing the file attributes failed");
  for (int64_t v = 0; v < a->n_vars; ++v) {
    const hsize_t dims[1] = {(hsize_t)n_values};
    Id space(H5Screate_simple(1, dims, nullptr), H5Sclose);
    const std::string name = "variables_" + std::to_string(v + 1);
    Id dset(H5Dcreate2(file.id, name.c_str(), H5T_IEEE_F64LE, space.id, H5P_DEFAULT, H5P_DEFAULT,
                       H5P_DEFAULT),
            H5Dclose);
    if (!space.ok() || !dset.ok() ||
        (n_values > 0 && H5Dwrite(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                  data + v * n_values) < 0) ||
        !write_attr_string(dset.id, "name", varnames[v] ? varnames[v] : ""))
      return h5_fail("writing " + name + " failed");
  }
  if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0) return h5_fail("flushing the file failed");
  return 0;
}

int atmo_h5_read_attributes(const char* filename, atmo_h5_attributes* a) {
  if (!filename || !a) return h5_fail("bad arguments");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  std::memset(a, 0, sizeof(*a));
  Id file(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return h5_fail(std::string("cannot open ") + filename);
  const bool ok = read_attr_int(file.id, "ndims", &a->ndims) &&
                  read_attr_string(file.id, "equations", a->equations, sizeof(a->equations)) &&
                  read_attr_int(file.id, "polydeg", &a->polydeg) &&
                  read_attr_int(file.id, "n_vars", &a->n_vars) &&
                  read_attr_int(file.id, "n_elements", &a->n_elements) &&
                  read_attr_string(file.id, "mesh_type", a->mesh_type, sizeof(a->mesh_type)) &&
                  read_attr_double(file.id, "time", &a->time) &&
                  read_attr_double(file.id, "dt", &a->dt) &&
                  read_attr_int(file.id, "timestep", &a->timestep);
  if (!ok) return h5_fail("missing or malformed file attributes");
  // mesh_file is absent from some writers' restart files
  if (H5Aexists(file.id, "mesh_file") > 0)
    read_attr_string(file.id, "mesh_file", a->mesh_file, sizeof(a->mesh_file));
  return 0;
}

int atmo_h5_read_variable(const char* filename, int32_t v, int64_t n_values, double* data,
                          char* name, int32_t name_capacity) {
  if (!filename || v < 1 || n_values < 0 || (n_values > 0 && !data))
    return h5_fail("bad arguments");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  Id file(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return h5_fail(std::string("cannot open ") + filename);
  const std::string dname = "variables_" + std::to_string(v);
  Id dset(H5Dopen2(file.id, dname.c_str(), H5P_DEFAULT), H5Dclose);
  if (!dset.ok()) return h5_fail("no dataset " + dname);
  Id space(H5Dget_space(dset.id), H5Sclose);
  if (!space.ok() || H5Sget_simple_extent_ndims(space.id) != 1 ||
      H5Sget_simple_extent_npoints(space.id) != (hssize_t)n_values)
    return h5_fail(dname + " does not hold " + std::to_string(n_values) + " values");
  if (n_values > 0 && H5Dread(dset.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return h5_fail("reading " + dname + " failed");
  if (name && name_capacity > 0) {
    name[0] = '\0';
    if (H5Aexists(dset.id, "name") > 0 && !read_attr_string(dset.id, "name", name, name_capacity))
      return h5_fail("malformed name attribute of " + dname);
  }
  return 0;
}

const char* atmo_h5_last_error(void) { return g_error.c_str(); }

}  // extern "C"
