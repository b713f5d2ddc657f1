// Reader of NumPy .npz archives (a zip of .npy files, stored or deflated), for networks saved by
// the Python side (WarmNet.save) and read by the commands.  Little-endian numeric arrays only.
#pragma once

#include <cstddef>
#include <map>
#include <string>
#include <vector>

namespace mrcsd {
namespace npz {

struct Array {
  std::string descr;            // numpy dtype string, e.g. "<f4", "<i8"
  std::vector<size_t> shape;
  bool fortran_order = false;
  std::vector<unsigned char> data;
  size_t count() const;
  // The values converted to double; throws for a dtype other than little-endian f4/f8/i4/i8/u1.
  std::vector<double> as_double() const;
};

// Arrays by name (the member name without ".npy"); throws std::runtime_error on a malformed file.
std::map<std::string, Array> load(const std::string& path);

}  // namespace npz
}  // namespace mrcsd
