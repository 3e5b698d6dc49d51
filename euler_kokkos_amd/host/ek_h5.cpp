// ek_h5.cpp -- HDF5 file access for the .ini driver's output and restart, built as its own
// shared library (libek_h5.so) against the system HDF5 and loaded on demand, so that
// nothing else depends on HDF5 being present.
//
// The file content is that of the reference's Save_HDF5 / Load_HDF5
// (src/utils/io/IO_HDF5.h:76-525, 1527-2143): one dataset per conservative variable at the
// root ("/rho", "/energy", ...), C order (z, y, x), chunked at the interior size with the
// shuffle and deflate filters (level 0 = stored), and root attributes "time step",
// "total time", "nx", "ny", "nz", "ghost zone included" (native int / double, scalar) and
// "creation date" (one variable-length string).
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>

namespace {

void set_err(char* err, int errlen, const std::string& m)
{
  if (err && errlen > 0) {
    std::strncpy(err, m.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = '\0';
  }
}

int put_int_attr(hid_t loc, const char* name, int v)
{
  const hid_t sp = H5Screate(H5S_SCALAR);
  const hid_t a = H5Acreate2(loc, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
  const herr_t s = a >= 0 ? H5Awrite(a, H5T_NATIVE_INT, &v) : -1;
  if (a >= 0)
    H5Aclose(a);
  H5Sclose(sp);
  return s < 0 ? -1 : 0;
}

// current_date (src/utils/io/IO_common.cpp:15-42): local time as "%Y-%m-%d %H:%M:%S %Z"
std::string current_date()
{
  std::time_t t = std::time(nullptr);
  std::tm tm = *std::localtime(&t);
  char buf[64];
  std::strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S %Z", &tm);
  return buf;
}

} // namespace

extern "C" {

// Write nvar datasets of shape dims[0..rank) (C order) named names[v] from data[v], plus the
// root attributes. chunk: chunk shape (rank entries). Returns 0 or -1 with a message.
int ek_h5_write(const char* filename, int rank, const unsigned long long* dims, const unsigned long long* chunk,
                int compression, int nvar, const char* const* names, const double* const* data, int time_step,
                double total_time, int nx, int ny, int nz, int ghost_included, char* err, int errlen)
{
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const hid_t f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) {
    set_err(err, errlen, std::string("cannot create ") + filename);
    return -1;
  }
  hsize_t d[3], c[3];
  for (int r = 0; r < rank; r++) {
    d[r] = (hsize_t)dims[r];
    c[r] = (hsize_t)chunk[r];
  }
  const hid_t space = H5Screate_simple(rank, d, nullptr);
  const hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  int rc = 0;
  if (H5Pset_chunk(dcpl, rank, c) < 0 || H5Pset_shuffle(dcpl) < 0 || H5Pset_deflate(dcpl, (unsigned)compression) < 0) {
    set_err(err, errlen, "cannot set the dataset creation properties");
    rc = -1;
  }
  for (int v = 0; v < nvar && rc == 0; v++) {
    const std::string name = std::string("/") + names[v];
    const hid_t ds = H5Dcreate2(f, name.c_str(), H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (ds < 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data[v]) < 0) {
      set_err(err, errlen, "cannot write dataset " + name);
      rc = -1;
    }
    if (ds >= 0)
      H5Dclose(ds);
  }
  if (rc == 0) {
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(f, "time step", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_INT, &time_step);
    H5Aclose(a);
    const hid_t b = H5Acreate2(f, "total time", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(b, H5T_NATIVE_DOUBLE, &total_time);
    H5Aclose(b);
    H5Sclose(sp);
    if (put_int_attr(f, "nx", nx) || put_int_attr(f, "ny", ny) || put_int_attr(f, "nz", nz) ||
        put_int_attr(f, "ghost zone included", ghost_included)) {
      set_err(err, errlen, "cannot write the root attributes");
      rc = -1;
    }
    const std::string date = current_date();
    const char* s = date.c_str();
    const hsize_t one = 1;
    const hid_t ty = H5Tcopy(H5T_C_S1);
    H5Tset_size(ty, H5T_VARIABLE);
    const hid_t root = H5Gopen2(f, "/", H5P_DEFAULT);
    const hid_t sp1 = H5Screate_simple(1, &one, nullptr);
    const hid_t at = H5Acreate2(root, "creation date", ty, sp1, H5P_DEFAULT, H5P_DEFAULT);
    if (at < 0 || H5Awrite(at, ty, &s) < 0) {
      set_err(err, errlen, "cannot write the creation date");
      rc = -1;
    }
    if (at >= 0)
      H5Aclose(at);
    H5Sclose(sp1);
    H5Gclose(root);
    H5Tclose(ty);
  }
  H5Pclose(dcpl);
  H5Sclose(space);
  H5Fflush(f, H5F_SCOPE_LOCAL);
  H5Fclose(f);
  return rc;
}

// Read datasets names[v] (each must have exactly the shape dims[0..rank), floating point)
// into out[v] as doubles, and the "time step" / "total time" root attributes.
int ek_h5_read(const char* filename, int rank, const unsigned long long* dims, int nvar, const char* const* names,
               double* const* out, int* time_step, double* total_time, char* err, int errlen)
{
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) {
    set_err(err, errlen, std::string("cannot open ") + filename);
    return -1;
  }
  int rc = 0;
  for (int v = 0; v < nvar && rc == 0; v++) {
    const std::string name = std::string("/") + names[v];
    const hid_t ds = H5Dopen2(f, name.c_str(), H5P_DEFAULT);
    if (ds < 0) {
      set_err(err, errlen, "no dataset " + name + " in " + filename);
      rc = -1;
      break;
    }
    const hid_t ty = H5Dget_type(ds);
    const hid_t sp = H5Dget_space(ds);
    hsize_t d[8] = {};
    const int r = H5Sget_simple_extent_ndims(sp);
    if (r == rank)
      H5Sget_simple_extent_dims(sp, d, nullptr);
    bool ok = r == rank && H5Tget_class(ty) == H5T_FLOAT;
    for (int k = 0; ok && k < rank; k++)
      ok = d[k] == (hsize_t)dims[k];
    if (!ok) {
      char m[256];
      std::snprintf(m, sizeof(m), "dataset %s: expected a floating-point array of rank %d and shape %llu x %llu x %llu",
                    name.c_str(), rank, dims[0], rank > 1 ? dims[1] : 1ull, rank > 2 ? dims[2] : 1ull);
      set_err(err, errlen, m);
      rc = -1;
    } else if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out[v]) < 0) {
      set_err(err, errlen, "cannot read dataset " + name);
      rc = -1;
    }
    H5Sclose(sp);
    H5Tclose(ty);
    H5Dclose(ds);
  }
  if (rc == 0) {
    const hid_t root = H5Gopen2(f, "/", H5P_DEFAULT);
    const hid_t a = H5Aopen(root, "time step", H5P_DEFAULT);
    const hid_t b = H5Aopen(root, "total time", H5P_DEFAULT);
    if (a < 0 || b < 0 || H5Aread(a, H5T_NATIVE_INT, time_step) < 0 || H5Aread(b, H5T_NATIVE_DOUBLE, total_time) < 0) {
      set_err(err, errlen, "cannot read the time step / total time attributes");
      rc = -1;
    }
    if (a >= 0)
      H5Aclose(a);
    if (b >= 0)
      H5Aclose(b);
    H5Gclose(root);
  }
  H5Fclose(f);
  return rc;
}

} // extern "C"
