// HDF5 files of the host layer through the HDF5 C library loaded at run time.
#include "h5.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "common.h"

namespace jams {
namespace h5 {
namespace {

// The C API's types and the constants used here (HDF5 1.10 and later:
// hid_t is 64 bits)
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr unsigned kAccRdonly = 0x0000u;  // H5F_ACC_RDONLY
constexpr unsigned kAccTrunc = 0x0002u;   // H5F_ACC_TRUNC
constexpr hid_t kDefault = 0;             // H5P_DEFAULT
constexpr hid_t kAll = 0;                 // H5S_ALL
constexpr int kScalar = 0;                // H5S_SCALAR
constexpr size_t kVariable = (size_t)-1;  // H5T_VARIABLE

struct Api {
  herr_t (*open)();
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  hid_t (*screate)(int);
  herr_t (*sclose)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  hid_t (*dget_space)(hid_t);
  herr_t (*dclose)(hid_t);
  hid_t (*pcreate)(hid_t);
  herr_t (*pset_chunk)(hid_t, int, const hsize_t *);
  herr_t (*pset_shuffle)(hid_t);
  herr_t (*pset_deflate)(hid_t, unsigned);
  herr_t (*pclose)(hid_t);
  hid_t (*tcopy)(hid_t);
  herr_t (*tset_size)(hid_t, size_t);
  herr_t (*tclose)(hid_t);
  hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
  herr_t (*awrite)(hid_t, hid_t, const void *);
  herr_t (*aclose)(hid_t);
  // the library's predefined identifiers (valid after H5open)
  const hid_t *native_double, *native_int, *c_s1, *dataset_create;
};

std::mutex g_mu;
bool g_tried = false;
Api g_api{};
std::string g_why;

template <class F>
bool bind(void *lib, const char *name, F &f) {
  f = reinterpret_cast<F>(dlsym(lib, name));
  return f != nullptr;
}

bool load_locked() {
  if (g_tried) return g_why.empty();
  g_tried = true;
  std::vector<std::string> names;
  if (const char *e = std::getenv("JAMS_HDF5_LIB")) names.push_back(e);
  names.push_back("libhdf5.so");
  names.push_back("/opt/conda/lib/libhdf5.so");
  void *lib = nullptr;
  std::string tried;
  for (const auto &n : names) {
    lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (lib) break;
    tried += (tried.empty() ? "" : ", ") + n;
  }
  if (!lib) {
    g_why = "no HDF5 C library could be loaded (tried " + tried + "; set JAMS_HDF5_LIB)";
    return false;
  }
  Api a{};
  bool ok = bind(lib, "H5open", a.open) && bind(lib, "H5Fcreate", a.fcreate) &&
            bind(lib, "H5Fopen", a.fopen) && bind(lib, "H5Fclose", a.fclose) &&
            bind(lib, "H5Screate_simple", a.screate_simple) && bind(lib, "H5Screate", a.screate) &&
            bind(lib, "H5Sclose", a.sclose) &&
            bind(lib, "H5Sget_simple_extent_ndims", a.sget_ndims) &&
            bind(lib, "H5Sget_simple_extent_dims", a.sget_dims) &&
            bind(lib, "H5Dcreate2", a.dcreate) && bind(lib, "H5Dopen2", a.dopen) &&
            bind(lib, "H5Dwrite", a.dwrite) && bind(lib, "H5Dread", a.dread) &&
            bind(lib, "H5Dget_space", a.dget_space) && bind(lib, "H5Dclose", a.dclose) &&
            bind(lib, "H5Pcreate", a.pcreate) && bind(lib, "H5Pset_chunk", a.pset_chunk) &&
            bind(lib, "H5Pset_shuffle", a.pset_shuffle) &&
            bind(lib, "H5Pset_deflate", a.pset_deflate) && bind(lib, "H5Pclose", a.pclose) &&
            bind(lib, "H5Tcopy", a.tcopy) && bind(lib, "H5Tset_size", a.tset_size) &&
            bind(lib, "H5Tclose", a.tclose) && bind(lib, "H5Acreate2", a.acreate) &&
            bind(lib, "H5Awrite", a.awrite) && bind(lib, "H5Aclose", a.aclose) &&
            bind(lib, "H5T_NATIVE_DOUBLE_g", a.native_double) &&
            bind(lib, "H5T_NATIVE_INT_g", a.native_int) && bind(lib, "H5T_C_S1_g", a.c_s1) &&
            bind(lib, "H5P_CLS_DATASET_CREATE_ID_g", a.dataset_create);
  if (!ok || a.open() < 0) {
    g_why = "the HDF5 C library lacks a function this build uses (HDF5 1.10 or later needed)";
    dlclose(lib);
    return false;
  }
  g_api = a;  // the library stays loaded until the process ends
  return true;
}

const Api &api() {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!load_locked()) throw RuntimeError("HDF5: " + g_why);
  return g_api;
}

void check(bool ok, const std::string &what) {
  if (!ok) throw RuntimeError("HDF5: " + what);
}

// closes an identifier on scope exit
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  ~Id() {
    if (id >= 0 && close) close(id);
  }
};

}  // namespace

bool available(std::string *why) {
  std::lock_guard<std::mutex> lock(g_mu);
  const bool ok = load_locked();
  if (!ok && why) *why = g_why;
  return ok;
}

void write_file(const std::string &file, const std::vector<Data> &sets, int chunk_rows,
                int level) {
  const Api &a = api();
  Id f{a.fcreate(file.c_str(), kAccTrunc, kDefault, kDefault), a.fclose};
  check(f.id >= 0, "failed to create " + file);
  for (const Data &d : sets) {
    check(!d.dims.empty() && d.dims.size() <= 2 && d.data, "bad dataset " + d.path);
    const int rank = (int)d.dims.size();
    hsize_t dims[2] = {(hsize_t)d.dims[0], rank > 1 ? (hsize_t)d.dims[1] : 0};
    Id space{a.screate_simple(rank, dims, nullptr), a.sclose};
    check(space.id >= 0, "dataspace of " + d.path);
    Id props{kDefault, nullptr};
    if (chunk_rows > 0 && dims[0] > 0) {
      props.id = a.pcreate(*a.dataset_create);
      props.close = a.pclose;
      check(props.id >= 0, "dataset properties of " + d.path);
      hsize_t chunk[2] = {std::min<hsize_t>((hsize_t)chunk_rows, dims[0]), 1};
      check(a.pset_chunk(props.id, rank, chunk) >= 0 && a.pset_shuffle(props.id) >= 0 &&
                a.pset_deflate(props.id, (unsigned)level) >= 0,
            "compression of " + d.path);
    }
    const hid_t type = d.integer ? *a.native_int : *a.native_double;
    Id set{a.dcreate(f.id, d.path.c_str(), type, space.id, kDefault, props.id, kDefault), a.dclose};
    check(set.id >= 0, "failed to create dataset " + d.path + " in " + file);
    check(a.dwrite(set.id, type, kAll, kAll, kDefault, d.data) >= 0,
          "failed to write dataset " + d.path + " in " + file);
    if (!d.units.empty()) {
      // a variable-length string, as the reference's writer stores it
      Id t{a.tcopy(*a.c_s1), a.tclose};
      check(t.id >= 0 && a.tset_size(t.id, kVariable) >= 0, "string type");
      Id s{a.screate(kScalar), a.sclose};
      Id at{a.acreate(set.id, "units", t.id, s.id, kDefault, kDefault), a.aclose};
      const char *v = d.units.c_str();
      check(s.id >= 0 && at.id >= 0 && a.awrite(at.id, t.id, &v) >= 0,
            "attribute units of " + d.path);
    }
  }
}

std::vector<double> read_doubles(const std::string &file, const std::string &path,
                                 std::vector<uint64_t> *dims) {
  const Api &a = api();
  Id f{a.fopen(file.c_str(), kAccRdonly, kDefault), a.fclose};
  check(f.id >= 0, "failed to open " + file);
  Id set{a.dopen(f.id, path.c_str(), kDefault), a.dclose};
  check(set.id >= 0, "no dataset " + path + " in " + file);
  Id space{a.dget_space(set.id), a.sclose};
  const int rank = space.id >= 0 ? a.sget_ndims(space.id) : -1;
  check(rank >= 1 && rank <= 8, "dataset " + path + " in " + file + " has no simple shape");
  hsize_t d[8];
  check(a.sget_dims(space.id, d, nullptr) == rank, "shape of " + path);
  size_t count = 1;
  if (dims) dims->clear();
  for (int k = 0; k < rank; ++k) {
    count *= (size_t)d[k];
    if (dims) dims->push_back((uint64_t)d[k]);
  }
  std::vector<double> out(count);
  if (count > 0)
    check(a.dread(set.id, *a.native_double, kAll, kAll, kDefault, out.data()) >= 0,
          "failed to read " + path + " from " + file);
  return out;
}

}  // namespace h5
}  // namespace jams
