// HDF5 files of the host layer: the spin snapshots and lattice file of the
// "hdf5" monitor (monitors/hdf5.cc:94-193) and the spin files that
// lattice.spins may name (helpers/load.h:13-19, :53-57).
//
// The HDF5 C library is resolved at run time (dlopen), so libjamshost.so
// itself links nothing beyond the C++ runtime and HIP: it is looked for as
// $JAMS_HDF5_LIB, then as libhdf5.so on the loader's search path, then in
// /opt/conda/lib.  Without it, writing or reading an HDF5 file is an error
// that names what was tried.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace jams {
namespace h5 {

// one dataset of a file to write
struct Data {
  std::string path;            // e.g. "/spins"
  bool integer = false;        // 32-bit signed integers, else doubles
  std::vector<uint64_t> dims;  // one or two dimensions, row-major
  const void *data = nullptr;
  std::string units;           // if set: a string attribute "units"
};

// true if the library can be loaded; otherwise false and why
bool available(std::string *why);

// Write a new file (an existing one is truncated) holding the datasets.
// chunk_rows > 0: two-dimensional datasets are stored in chunks of
// (min(chunk_rows, rows), 1) compressed with deflate at level `level`.
void write_file(const std::string &file, const std::vector<Data> &sets, int chunk_rows = 0,
                int level = 6);

// The double dataset `path` of `file`, row-major, and its dimensions.
std::vector<double> read_doubles(const std::string &file, const std::string &path,
                                 std::vector<uint64_t> *dims);

}  // namespace h5
}  // namespace jams
