// ptycxi: CXI-style result files (ptychography/display.py:21-117) through the
// HDF5 C library, in the layout h5py gives the same numpy arrays.  Host code
// only; see include/ptycxi.h for the interface.
#include <hdf5.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ptycxi.h"
#include "ptyhip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Closes an HDF5 handle of any kind when it goes out of scope.
struct Handle {
  hid_t id = H5I_INVALID_HID;
  explicit Handle(hid_t h = H5I_INVALID_HID) : id(h) {}
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  ~Handle() {
    if (id >= 0) {
      switch (H5Iget_type(id)) {
        case H5I_FILE: H5Fclose(id); break;
        case H5I_GROUP: H5Gclose(id); break;
        case H5I_DATASET: H5Dclose(id); break;
        case H5I_DATASPACE: H5Sclose(id); break;
        case H5I_DATATYPE: H5Tclose(id); break;
        case H5I_GENPROP_LST: H5Pclose(id); break;
        default: break;
      }
    }
  }
  bool ok() const { return id >= 0; }
};

// Silences HDF5's own error printing for the lifetime of the object (failures
// are reported through the return codes and pty_cxi_last_error()).
struct Quiet {
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

size_t elem_size(int dtype) {
  switch (dtype) {
    case PTY_CXI_F32: case PTY_CXI_I32: return 4;
    case PTY_CXI_F64: case PTY_CXI_I64: case PTY_CXI_C64: return 8;
    case PTY_CXI_C128: return 16;
    case PTY_CXI_BOOL: return 1;
    default: return 0;
  }
}

// h5py's complex layout: a compound of two floats named "r" and "i"
hid_t complex_type(hid_t flt, size_t half) {
  const hid_t t = H5Tcreate(H5T_COMPOUND, 2 * half);
  if (t < 0) return t;
  H5Tinsert(t, "r", 0, flt);
  H5Tinsert(t, "i", half, flt);
  return t;
}

// h5py's bool layout: an int8 enum FALSE = 0, TRUE = 1
hid_t bool_type() {
  const hid_t t = H5Tenum_create(H5T_NATIVE_INT8);
  if (t < 0) return t;
  int8_t v = 0;
  H5Tenum_insert(t, "FALSE", &v);
  v = 1;
  H5Tenum_insert(t, "TRUE", &v);
  return t;
}

// type written to the file for a layout
hid_t file_type(int dtype) {
  switch (dtype) {
    case PTY_CXI_F32: return H5Tcopy(H5T_IEEE_F32LE);
    case PTY_CXI_F64: return H5Tcopy(H5T_IEEE_F64LE);
    case PTY_CXI_C64: return complex_type(H5T_IEEE_F32LE, 4);
    case PTY_CXI_C128: return complex_type(H5T_IEEE_F64LE, 8);
    case PTY_CXI_I32: return H5Tcopy(H5T_STD_I32LE);
    case PTY_CXI_I64: return H5Tcopy(H5T_STD_I64LE);
    case PTY_CXI_BOOL: return bool_type();
    default: return H5I_INVALID_HID;
  }
}

// the same layout in host memory
hid_t mem_type(int dtype) {
  switch (dtype) {
    case PTY_CXI_F32: return H5Tcopy(H5T_NATIVE_FLOAT);
    case PTY_CXI_F64: return H5Tcopy(H5T_NATIVE_DOUBLE);
    case PTY_CXI_C64: return complex_type(H5T_NATIVE_FLOAT, 4);
    case PTY_CXI_C128: return complex_type(H5T_NATIVE_DOUBLE, 8);
    case PTY_CXI_I32: return H5Tcopy(H5T_NATIVE_INT32);
    case PTY_CXI_I64: return H5Tcopy(H5T_NATIVE_INT64);
    case PTY_CXI_BOOL: return bool_type();
    default: return H5I_INVALID_HID;
  }
}

// layout of a stored type, or 0 if none of the above
int classify(hid_t t) {
  const H5T_class_t cls = H5Tget_class(t);
  const size_t sz = H5Tget_size(t);
  if (cls == H5T_FLOAT) return sz == 4 ? PTY_CXI_F32 : sz == 8 ? PTY_CXI_F64 : 0;
  if (cls == H5T_INTEGER) return sz <= 4 ? PTY_CXI_I32 : sz == 8 ? PTY_CXI_I64 : 0;
  if (cls == H5T_ENUM) return H5Tget_nmembers(t) == 2 && sz == 1 ? PTY_CXI_BOOL : 0;
  if (cls == H5T_COMPOUND && H5Tget_nmembers(t) == 2) {
    char* n0 = H5Tget_member_name(t, 0);
    char* n1 = H5Tget_member_name(t, 1);
    const bool ri = n0 && n1 && !std::strcmp(n0, "r") && !std::strcmp(n1, "i");
    if (n0) H5free_memory(n0);
    if (n1) H5free_memory(n1);
    if (!ri) return 0;
    Handle m0(H5Tget_member_type(t, 0));
    if (H5Tget_class(m0.id) != H5T_FLOAT) return 0;
    return H5Tget_size(m0.id) == 4 ? PTY_CXI_C64 : H5Tget_size(m0.id) == 8 ? PTY_CXI_C128 : 0;
  }
  return 0;
}

bool exists(hid_t file, const std::string& name) {
  // every intermediate link must exist before H5Lexists may be asked about the next
  size_t pos = 0;
  while (true) {
    pos = name.find('/', pos + 1);
    const std::string part = name.substr(0, pos);
    if (H5Lexists(file, part.c_str(), H5P_DEFAULT) <= 0) return false;
    if (pos == std::string::npos) return true;
  }
}

}  // namespace

extern "C" {

const char* pty_cxi_last_error(void) { return g_err.c_str(); }

int pty_cxi_create(const char* path) {
  if (!path) return fail(PTY_EINVAL, "path is NULL");
  Quiet q;
  Handle f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT));
  if (!f.ok()) return fail(PTY_EHIP, std::string("cannot create ") + path);
  return PTY_OK;
}

int pty_cxi_write(const char* path, const char* name, int32_t dtype, int32_t ndim,
                  const int64_t* dims, const void* data, int32_t gzip_level) {
  if (!path || !name || !data || (ndim > 0 && !dims)) return fail(PTY_EINVAL, "NULL argument");
  if (ndim < 0 || ndim > PTY_CXI_MAXDIM) return fail(PTY_EINVAL, "ndim must be 0..8");
  if (!elem_size(dtype)) return fail(PTY_EINVAL, "unknown dtype");
  if (gzip_level < 0 || gzip_level > 9) return fail(PTY_EINVAL, "gzip level must be 0..9");
  std::vector<hsize_t> d(ndim);
  size_t row = elem_size(dtype);
  for (int k = 0; k < ndim; ++k) {
    if (dims[k] < 0) return fail(PTY_EINVAL, "negative dimension");
    d[k] = (hsize_t)dims[k];
    if (k > 0) row *= (size_t)dims[k];
  }
  Quiet q;
  Handle f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT));
  if (!f.ok()) return fail(PTY_EHIP, std::string("cannot open ") + path + " for writing");
  if (exists(f.id, name)) return fail(PTY_EINVAL, std::string(name) + " already exists");
  Handle space(ndim == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(ndim, d.data(), nullptr));
  Handle ft(file_type(dtype)), mt(mem_type(dtype));
  Handle lcpl(H5Pcreate(H5P_LINK_CREATE));
  Handle dcpl(H5Pcreate(H5P_DATASET_CREATE));
  if (!space.ok() || !ft.ok() || !mt.ok() || !lcpl.ok() || !dcpl.ok())
    return fail(PTY_EHIP, "HDF5 type / space / property set-up failed");
  H5Pset_create_intermediate_group(lcpl.id, 1);
  bool any_zero = false;
  for (int k = 0; k < ndim; ++k) any_zero |= d[k] == 0;
  if (gzip_level > 0 && ndim > 0 && !any_zero) {
    // chunks of whole rows along the first axis, about 1 MiB each
    std::vector<hsize_t> c(d);
    const size_t per = row ? (size_t)1 << 20 : 1;
    c[0] = std::max<hsize_t>(1, std::min<hsize_t>(d[0], row ? per / row : d[0]));
    if (H5Pset_chunk(dcpl.id, ndim, c.data()) < 0 || H5Pset_deflate(dcpl.id, gzip_level) < 0)
      return fail(PTY_EHIP, "chunked deflate layout not available");
  }
  Handle ds(H5Dcreate2(f.id, name, ft.id, space.id, lcpl.id, dcpl.id, H5P_DEFAULT));
  if (!ds.ok()) return fail(PTY_EHIP, std::string("cannot create dataset ") + name);
  if (!any_zero && H5Dwrite(ds.id, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(PTY_EHIP, std::string("writing ") + name + " failed");
  return PTY_OK;
}

int pty_cxi_info(const char* path, const char* name, int32_t* dtype, int32_t* ndim, int64_t* dims) {
  if (!path || !name || !dtype || !ndim || !dims) return fail(PTY_EINVAL, "NULL argument");
  Quiet q;
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
  if (!f.ok()) return fail(PTY_EHIP, std::string("cannot open ") + path);
  if (!exists(f.id, name)) return fail(PTY_CXI_ENOENT, std::string("no dataset ") + name);
  Handle ds(H5Dopen2(f.id, name, H5P_DEFAULT));
  if (!ds.ok()) return fail(PTY_EHIP, std::string("cannot open dataset ") + name);
  Handle t(H5Dget_type(ds.id)), space(H5Dget_space(ds.id));
  const int cls = classify(t.id);
  if (!cls) return fail(PTY_EINVAL, std::string(name) + " has a type other than float, integer, "
                                                       "bool or complex");
  const int nd = H5Sget_simple_extent_ndims(space.id);
  if (nd < 0 || nd > PTY_CXI_MAXDIM) return fail(PTY_EINVAL, std::string(name) + ": rank above 8");
  std::vector<hsize_t> d(nd > 0 ? nd : 1);
  H5Sget_simple_extent_dims(space.id, d.data(), nullptr);
  *dtype = cls;
  *ndim = nd;
  for (int k = 0; k < nd; ++k) dims[k] = (int64_t)d[k];
  return PTY_OK;
}

int pty_cxi_read(const char* path, const char* name, int32_t dtype, void* data, int64_t max_rows) {
  if (!path || !name || !data) return fail(PTY_EINVAL, "NULL argument");
  if (!elem_size(dtype)) return fail(PTY_EINVAL, "unknown dtype");
  Quiet q;
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT));
  if (!f.ok()) return fail(PTY_EHIP, std::string("cannot open ") + path);
  if (!exists(f.id, name)) return fail(PTY_CXI_ENOENT, std::string("no dataset ") + name);
  Handle ds(H5Dopen2(f.id, name, H5P_DEFAULT));
  if (!ds.ok()) return fail(PTY_EHIP, std::string("cannot open dataset ") + name);
  Handle t(H5Dget_type(ds.id)), space(H5Dget_space(ds.id));
  const int cls = classify(t.id);
  const bool cplx_s = cls == PTY_CXI_C64 || cls == PTY_CXI_C128;
  const bool cplx_d = dtype == PTY_CXI_C64 || dtype == PTY_CXI_C128;
  if (!cls || cplx_s != cplx_d || (cls == PTY_CXI_BOOL) != (dtype == PTY_CXI_BOOL))
    return fail(PTY_EINVAL, std::string(name) + ": stored type does not convert to the one asked for");
  Handle mt(mem_type(dtype));
  const int nd = H5Sget_simple_extent_ndims(space.id);
  std::vector<hsize_t> d(nd > 0 ? nd : 1);
  H5Sget_simple_extent_dims(space.id, d.data(), nullptr);
  bool any_zero = false;
  for (int k = 0; k < nd; ++k) any_zero |= d[k] == 0;
  if (nd > 0 && max_rows >= 0 && (hsize_t)max_rows < d[0]) d[0] = (hsize_t)max_rows;
  if (any_zero || (nd > 0 && d[0] == 0)) return PTY_OK;
  herr_t e;
  if (nd > 0 && max_rows >= 0) {
    std::vector<hsize_t> start(nd, 0);
    Handle mspace(H5Screate_simple(nd, d.data(), nullptr));
    H5Sselect_hyperslab(space.id, H5S_SELECT_SET, start.data(), nullptr, d.data(), nullptr);
    e = H5Dread(ds.id, mt.id, mspace.id, space.id, H5P_DEFAULT, data);
  } else {
    e = H5Dread(ds.id, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  }
  if (e < 0) return fail(PTY_EHIP, std::string("reading ") + name + " failed");
  return PTY_OK;
}

}  // extern "C"
