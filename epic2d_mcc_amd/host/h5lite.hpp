// A minimal, self-contained HDF5 file writer: enough of the HDF5 file
// format (version 2 superblock, version 2 object headers, compact link
// storage) for the layout of the reference's result.h5 (output.cpp:4-315):
// nested groups, contiguous two-dimensional FP64 datasets, and scalar
// attributes (32-bit signed integer, FP64, fixed-length ASCII string).
//
// Dataset values go to the file as they are created (after a 48-byte slot
// reserved for the superblock); the object headers, which hold the links
// between groups and the datasets' addresses, are written when the file is
// closed, then the superblock.  Every piece of metadata carries the lookup3
// checksum the format requires.  Readers: any HDF5 library of version 1.8
// or later.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace epic2d {

// Bob Jenkins' lookup3 hash, byte-wise hashlittle (the checksum of HDF5
// version 2 metadata structures).
uint32_t lookup3(const void* data, size_t length, uint32_t initval);

class H5Lite {
 public:
  static constexpr int ROOT = 0;

  explicit H5Lite(const std::string& path);  // creates or truncates
  ~H5Lite();
  H5Lite(const H5Lite&) = delete;
  H5Lite& operator=(const H5Lite&) = delete;

  // The child group `name` of `parent` (created when absent); -1 if `name`
  // names a dataset.
  int group(int parent, const std::string& name);
  // Child object of `parent` called `name`, or -1.
  int find(int parent, const std::string& name) const;
  // A new FP64 dataset of the given dimensions (rank 1..4), its values in
  // row-major order; -1 if the name is taken.
  int dataset_f64(int parent, const std::string& name, const std::vector<uint64_t>& dims, const double* data);
  // Scalar attributes of a group or dataset; false if the name is taken.
  bool attr_i32(int obj, const std::string& name, int32_t v);
  bool attr_f64(int obj, const std::string& name, double v);
  // A fixed-length, NUL-terminated ASCII string of `size` bytes holding the
  // first bytes of `s` (the reference's PredType::C_S1 is size 1).
  bool attr_str(int obj, const std::string& name, const std::string& s, uint32_t size);

  // Writes the object headers and the superblock; false on an I/O error.
  bool close();
  bool ok() const { return f_ != nullptr && !failed_; }

 private:
  struct Attr {
    std::string name;
    std::vector<uint8_t> type, space, data;
  };
  struct Obj {
    bool is_group = true;
    std::string name;
    int parent = -1;
    std::vector<int> kids;
    std::vector<Attr> attrs;
    std::vector<uint64_t> dims;
    uint64_t data_addr = 0, data_size = 0;
  };
  bool add_attr(int obj, Attr a);
  std::vector<uint8_t> header(const Obj& o, const std::vector<uint64_t>& addr) const;

  FILE* f_ = nullptr;
  bool failed_ = false;
  uint64_t eof_ = 0;
  std::vector<Obj> objs_;
};

}  // namespace epic2d
