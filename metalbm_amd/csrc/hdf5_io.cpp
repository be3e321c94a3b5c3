// hdf5_io.cpp -- field and distribution files in the reference's HDF5 layout.
//
// The reference writes one FP64 dataset per field component, in the order of
// FieldList::writeFields (FieldList.h:65-86): "density", "velocityX",
// "velocityY"[, "velocityZ"], "alpha" (when written), "forceX".. (when
// written), and one per population ("distribution0" .. "distribution{Q-1}"),
// each of dimensions gSD::pLength() (Writer.h:309-360, 401-440): the global
// extents with the last one padded to 2 (n/2 + 1) for FFTW's in-place real
// transforms (Domain.h:53-57, MathVector.h:331-344).  Restart reads the
// distribution datasets back (Reader.h:106-158, Initialize.h:119-145).  Values here are
// the unshifted FP64 populations and the stored FP64 fields; the padding is
// written as zero and ignored on reading.  The XDMF sidecar follows
// FieldWriter<XDMF> (Writer.h:451-560).
//
// libhdf5 (1.10 or later, where hid_t is 64-bit) is loaded on first use with
// dlopen -- the library itself has no link-time HDF5 dependency.  Search
// order: $METALBM_HDF5_LIBRARY, libhdf5.so, libhdf5.so.103, then
// /opt/conda/lib.  Built on the public entry points only.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "hdf5_layout.hpp"

namespace {

using lbm::Grid;
using lbm::from_padded;
using lbm::hsize_t;
using lbm::set_extents;
using lbm::to_padded;

typedef int64_t hid_t;
typedef int herr_t;
constexpr hid_t kDefault = 0;           // H5P_DEFAULT, H5E_DEFAULT
constexpr unsigned kAccTrunc = 0x0002u;  // H5F_ACC_TRUNC
constexpr unsigned kAccRdonly = 0x0000u;  // H5F_ACC_RDONLY

struct H5 {
  void* lib = nullptr;
  hid_t native_double = -1;
  herr_t (*open)();
  herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*);
  herr_t (*eset_auto)(hid_t, void*, void*);
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char*, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
  hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char*, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  hid_t (*dget_space)(hid_t);
  herr_t (*dclose)(hid_t);
  std::string error;
};

template <class F>
bool bind(void* lib, const char* name, F& fn) {
  fn = reinterpret_cast<F>(dlsym(lib, name));
  return fn != nullptr;
}

const H5& hdf5() {
  static const H5 h = [] {
    H5 x;
    std::vector<std::string> names;
    if (const char* e = std::getenv("METALBM_HDF5_LIBRARY")) names.push_back(e);
    for (const char* n : {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
                          "/opt/conda/lib/libhdf5.so.103"})
      names.push_back(n);
    for (const std::string& n : names)
      if ((x.lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!x.lib) {
      x.error = "HDF5 library not found (set METALBM_HDF5_LIBRARY)";
      return x;
    }
    bool ok = bind(x.lib, "H5open", x.open) && bind(x.lib, "H5get_libversion", x.get_libversion) &&
              bind(x.lib, "H5Eset_auto2", x.eset_auto) && bind(x.lib, "H5Fcreate", x.fcreate) &&
              bind(x.lib, "H5Fopen", x.fopen) && bind(x.lib, "H5Fclose", x.fclose) &&
              bind(x.lib, "H5Screate_simple", x.screate_simple) &&
              bind(x.lib, "H5Sclose", x.sclose) &&
              bind(x.lib, "H5Sget_simple_extent_ndims", x.sget_ndims) &&
              bind(x.lib, "H5Sget_simple_extent_dims", x.sget_dims) &&
              bind(x.lib, "H5Dcreate2", x.dcreate) && bind(x.lib, "H5Dopen2", x.dopen) &&
              bind(x.lib, "H5Dwrite", x.dwrite) && bind(x.lib, "H5Dread", x.dread) &&
              bind(x.lib, "H5Dget_space", x.dget_space) && bind(x.lib, "H5Dclose", x.dclose);
    unsigned maj = 0, min = 0, rel = 0;
    if (ok) ok = x.open() >= 0 && x.get_libversion(&maj, &min, &rel) >= 0;
    if (ok && (maj < 1 || (maj == 1 && min < 10))) {
      x.error = "HDF5 1.10 or later is required";
      return x;
    }
    const hid_t* nd = ok ? static_cast<const hid_t*>(dlsym(x.lib, "H5T_NATIVE_DOUBLE_g")) : nullptr;
    if (!ok || !nd) {
      x.error = "the HDF5 library lacks a required symbol";
      return x;
    }
    x.native_double = *nd;
    x.eset_auto(kDefault, nullptr, nullptr);  // errors come back as return codes
    return x;
  }();
  return h;
}

int grid_of(lbm_handle* h, Grid& g) {
  lbm_config c;
  if (int rc = lbm_get_config(h, &c)) return rc;
  int64_t N = 0;
  int Q = 0, D = 0;
  if (int rc = lbm_sizes(h, &N, &Q, &D)) return rc;
  set_extents(g, D, Q, c.nx, c.ny, c.nz);
  return LBM_OK;
}

class File {
 public:
  File(const H5& x, hid_t id) : x_(x), id_(id) {}
  ~File() {
    if (id_ >= 0) x_.fclose(id_);
  }
  hid_t id() const { return id_; }

 private:
  const H5& x_;
  hid_t id_;
};

bool write_dataset(const H5& x, hid_t file, const Grid& g, const std::string& name,
                   const double* dense) {
  std::vector<double> buf;
  to_padded(g, dense, buf);
  const hid_t space = x.screate_simple(g.D, g.dims, nullptr);
  if (space < 0) return false;
  const hid_t set = x.dcreate(file, name.c_str(), x.native_double, space, kDefault, kDefault,
                              kDefault);
  bool ok = set >= 0;
  if (ok) ok = x.dwrite(set, x.native_double, space, space, kDefault, buf.data()) >= 0;
  if (set >= 0) x.dclose(set);
  x.sclose(space);
  return ok;
}

// Reads one dataset of the padded dimensions; false (with why) otherwise.
bool read_dataset(const H5& x, hid_t file, const Grid& g, const std::string& name,
                  double* dense, std::string& why) {
  const hid_t set = x.dopen(file, name.c_str(), kDefault);
  if (set < 0) {
    why = "dataset " + name + " not found";
    return false;
  }
  const hid_t space = x.dget_space(set);
  hsize_t dims[8] = {0};
  const int nd = space >= 0 ? x.sget_ndims(space) : -1;
  bool ok = nd == g.D && nd <= 8 && x.sget_dims(space, dims, nullptr) == nd;
  for (int d = 0; ok && d < g.D; ++d) ok = dims[d] == g.dims[d];
  if (!ok) why = "dataset " + name + " does not have the padded extents of this grid";
  std::vector<double> buf(g.padded);
  if (ok) {
    ok = x.dread(set, x.native_double, space, space, kDefault, buf.data()) >= 0;
    if (!ok) why = "reading " + name + " failed";
  }
  if (ok) from_padded(g, buf, dense);
  if (space >= 0) x.sclose(space);
  x.dclose(set);
  return ok;
}

const char* kAxis[3] = {"X", "Y", "Z"};

void write_xdmf(const Grid& g, const std::string& path, const std::string& h5name,
                const std::vector<std::string>& names) {
  std::ofstream o(path);
  auto dims = [&] {
    std::string s = std::to_string(g.dims[0]);
    for (int d = 1; d < g.D; ++d) s += " " + std::to_string(g.n[d]);
    return s;
  };
  o << "<?xml version=\"1.0\"?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf>\n<Domain>\n"
    << "<Grid Name=\"grid\" GridType=\"Uniform\">\n"
    << "<Topology TopologyType=\"" << g.D << "DCoRectMesh\" Dimensions=\"" << dims()
    << "\" />\n<Geometry GeometryType=\"Origin_Dx" << (g.D >= 2 ? "Dy" : "")
    << (g.D == 3 ? "Dz" : "") << "\">\n<DataItem Dimensions=\"" << g.D
    << "\" NumberType=\"Integer\" Format=\"XML\">0";
  for (int d = 1; d < g.D; ++d) o << " 0";
  o << "</DataItem>\n<DataItem Dimensions=\"" << g.D
    << "\" NumberType=\"Integer\" Format=\"XML\">1";
  for (int d = 1; d < g.D; ++d) o << " 1";
  o << "</DataItem>\n</Geometry>\n";
  for (const std::string& n : names)
    o << "<Attribute Name=\"" << n << "\" AttributeType=\"Scalar\" Center=\"Node\">\n"
      << "<DataItem Dimensions=\"" << dims()
      << "\" NumberType=\"Double\" Precision=\"8\" Format=\"HDF\">\n"
      << h5name << ":/" << n << "\n</DataItem>\n</Attribute>\n";
  o << "</Grid>\n</Domain>\n</Xdmf>\n";
}

std::string base_name(const std::string& p) {
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? p : p.substr(s + 1);
}

}  // namespace

extern "C" {

int lbm_write_fields_hdf5(lbm_handle* h, const char* path, const char* xdmf_path) {
  if (!h || !path) return h ? lbm::record_error(h, LBM_E_INVALID, "null path") : LBM_E_INVALID;
  const H5& x = hdf5();
  if (!x.error.empty()) return lbm::record_error(h, LBM_E_STATE, x.error);
  Grid g;
  if (int rc = grid_of(h, g)) return rc;
  std::vector<double> rho(g.N), u((size_t)g.D * g.N), alpha(g.N), F((size_t)g.D * g.N);
  if (int rc = lbm_get_fields(h, rho.data(), u.data(), alpha.data(), F.data())) return rc;
  File f(x, x.fcreate(path, kAccTrunc, kDefault, kDefault));
  if (f.id() < 0) return lbm::record_error(h, LBM_E_INVALID, std::string("cannot create ") + path);
  int w_alpha = 1, w_force = 1, w_vort = 1;
  lbm::field_output(h, &w_alpha, &w_force, &w_vort);
  // vorticity as the reference stores it: curl of u_h by Curl::executeSpace +
  // normalize (Routine.h:129-132), one component in 2D, three in 3D
  const int C = 2 * g.D - 3;
  std::vector<double> omega;
  if (w_vort) {
    omega.resize((size_t)C * g.N);
    if (int rc = lbm_vorticity(h, LBM_SPECTRAL_REFERENCE, omega.data(), nullptr)) return rc;
  }
  std::vector<std::string> names;
  auto put = [&](const std::string& name, const double* data) {
    names.push_back(name);
    return write_dataset(x, f.id(), g, name, data);
  };
  // FieldList::writeFields order (FieldList.h:65-86); a field of several
  // components gets the axis suffix per component (Writer.h:317-320)
  bool ok = put("density", rho.data());
  for (int d = 0; ok && d < g.D; ++d)
    ok = put(std::string("velocity") + kAxis[d], u.data() + (size_t)d * g.N);
  if (ok && w_alpha) ok = put("alpha", alpha.data());
  for (int d = 0; ok && w_force && d < g.D; ++d)
    ok = put(std::string("force") + kAxis[d], F.data() + (size_t)d * g.N);
  for (int k = 0; ok && w_vort && k < C; ++k)
    ok = put(C == 1 ? std::string("vorticity") : std::string("vorticity") + kAxis[k],
             omega.data() + (size_t)k * g.N);
  if (!ok) return lbm::record_error(h, LBM_E_INVALID, std::string("writing ") + path + " failed");
  if (xdmf_path) write_xdmf(g, xdmf_path, base_name(path), names);
  return LBM_OK;
}

int lbm_write_distribution_hdf5(lbm_handle* h, const char* path) {
  if (!h || !path) return h ? lbm::record_error(h, LBM_E_INVALID, "null path") : LBM_E_INVALID;
  const H5& x = hdf5();
  if (!x.error.empty()) return lbm::record_error(h, LBM_E_STATE, x.error);
  Grid g;
  if (int rc = grid_of(h, g)) return rc;
  std::vector<double> f((size_t)g.Q * g.N);
  if (int rc = lbm_get_populations(h, f.data())) return rc;
  File file(x, x.fcreate(path, kAccTrunc, kDefault, kDefault));
  if (file.id() < 0)
    return lbm::record_error(h, LBM_E_INVALID, std::string("cannot create ") + path);
  for (int i = 0; i < g.Q; ++i)
    if (!write_dataset(x, file.id(), g, "distribution" + std::to_string(i),
                       f.data() + (size_t)i * g.N))
      return lbm::record_error(h, LBM_E_INVALID, std::string("writing ") + path + " failed");
  return LBM_OK;
}

int lbm_read_distribution_hdf5(lbm_handle* h, const char* path) {
  if (!h || !path) return h ? lbm::record_error(h, LBM_E_INVALID, "null path") : LBM_E_INVALID;
  const H5& x = hdf5();
  if (!x.error.empty()) return lbm::record_error(h, LBM_E_STATE, x.error);
  Grid g;
  if (int rc = grid_of(h, g)) return rc;
  File file(x, x.fopen(path, kAccRdonly, kDefault));
  if (file.id() < 0) return lbm::record_error(h, LBM_E_INVALID, std::string("cannot open ") + path);
  std::vector<double> f((size_t)g.Q * g.N);
  std::string why;
  for (int i = 0; i < g.Q; ++i)
    if (!read_dataset(x, file.id(), g, "distribution" + std::to_string(i),
                      f.data() + (size_t)i * g.N, why))
      return lbm::record_error(h, LBM_E_INVALID, why);
  if (int rc = lbm_set_populations(h, f.data())) return rc;
  // a restart starts the entropic solve from alpha = 2 as a fresh run does
  // (initAlpha, Initialize.h:81-88)
  return lbm::reset_alpha(h);
}

}  // extern "C"
