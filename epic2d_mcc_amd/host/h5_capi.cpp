// C-ABI of the result.h5 writer (include/epic2d_h5.h).
#include <new>

#include "epic2d_h5.h"
#include "h5lite.hpp"

struct epic_h5 {
  epic2d::H5Lite w;
  explicit epic_h5(const char* path) : w(path) {}
};

extern "C" {

int epic_h5_create(const char* path, epic_h5** out) {
  if (!path || !out) return EPIC_ERR_ARG;
  *out = nullptr;
  epic_h5* h = new (std::nothrow) epic_h5(path);
  if (!h) return EPIC_ERR_ARG;
  if (!h->w.ok()) {
    delete h;
    return EPIC_ERR_ARG;
  }
  *out = h;
  return EPIC_OK;
}

int epic_h5_group(epic_h5* h, int parent, const char* name, int* out) {
  if (!h || !name || !out) return EPIC_ERR_ARG;
  *out = h->w.group(parent, name);
  return *out < 0 ? EPIC_ERR_ARG : EPIC_OK;
}

int epic_h5_dataset_f64(epic_h5* h, int parent, const char* name, int rank, const uint64_t* dims,
                        const double* data, int* out) {
  if (!h || !name || !out || !dims || rank < 1 || rank > 4) return EPIC_ERR_ARG;
  std::vector<uint64_t> d(dims, dims + rank);
  uint64_t n = 1;
  for (uint64_t v : d) n *= v;
  if (n && !data) return EPIC_ERR_ARG;
  *out = h->w.dataset_f64(parent, name, d, data);
  return *out < 0 ? EPIC_ERR_ARG : EPIC_OK;
}

int epic_h5_attr_i32(epic_h5* h, int obj, const char* name, int32_t v) {
  if (!h || !name) return EPIC_ERR_ARG;
  return h->w.attr_i32(obj, name, v) ? EPIC_OK : EPIC_ERR_ARG;
}

int epic_h5_attr_f64(epic_h5* h, int obj, const char* name, double v) {
  if (!h || !name) return EPIC_ERR_ARG;
  return h->w.attr_f64(obj, name, v) ? EPIC_OK : EPIC_ERR_ARG;
}

int epic_h5_attr_str(epic_h5* h, int obj, const char* name, const char* s, uint32_t size) {
  if (!h || !name || !s) return EPIC_ERR_ARG;
  return h->w.attr_str(obj, name, s, size) ? EPIC_OK : EPIC_ERR_ARG;
}

int epic_h5_close(epic_h5* h) {
  if (!h) return EPIC_ERR_ARG;
  const bool good = h->w.close();
  delete h;
  return good ? EPIC_OK : EPIC_ERR_ARG;
}

uint32_t epic_h5_lookup3(const void* data, size_t length, uint32_t initval) {
  return epic2d::lookup3(data, length, initval);
}

}  // extern "C"
