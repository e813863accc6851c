/* epic2d_h5.h — the result.h5 writer of the host library (libepic2d_host).
 *
 * Replaces the HDF5 C++ API calls of the reference's Output class
 * (output.cpp:4-315: H5File, Group::createGroup, createDataSet + write of
 * NATIVE_DOUBLE arrays, createAttribute + write of NATIVE_INT, NATIVE_DOUBLE
 * and C_S1 scalars) with a self-contained writer of the HDF5 file format
 * (version 2 superblock and object headers, compact links, contiguous
 * datasets), so the build needs no HDF5 library.  The host class Output
 * (host/epic_host.hpp) writes the reference's layout through it; these entry
 * points expose the writer itself.
 *
 * Objects are numbered: 0 is the root group; every group and dataset call
 * returns the new object's number.  Names are single path components.
 * Dataset values are written to the file at once; the group structure and
 * attributes when epic_h5_close is called, which also frees the handle.
 */
#ifndef EPIC2D_H5_H_
#define EPIC2D_H5_H_

#include <stddef.h>
#include <stdint.h>

#include "epic2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct epic_h5 epic_h5;

/* Creates (truncates) the file.  EPIC_ERR_ARG if it cannot be opened. */
int epic_h5_create(const char* path, epic_h5** out);
/* The child group `name` of group `parent`, created if absent. */
int epic_h5_group(epic_h5* h, int parent, const char* name, int* out);
/* A new FP64 dataset, rank 1..4, values row-major (C order). */
int epic_h5_dataset_f64(epic_h5* h, int parent, const char* name, int rank, const uint64_t* dims,
                        const double* data, int* out);
/* Scalar attributes of a group or dataset (EPIC_ERR_ARG if the name is taken). */
int epic_h5_attr_i32(epic_h5* h, int obj, const char* name, int32_t v);
int epic_h5_attr_f64(epic_h5* h, int obj, const char* name, double v);
/* Fixed-length NUL-terminated ASCII string of `size` bytes (the first bytes of s). */
int epic_h5_attr_str(epic_h5* h, int obj, const char* name, const char* s, uint32_t size);
/* Writes the metadata and the superblock, closes the file and frees h. */
int epic_h5_close(epic_h5* h);
/* The format's metadata checksum: Bob Jenkins' lookup3 hashlittle. */
uint32_t epic_h5_lookup3(const void* data, size_t length, uint32_t initval);

#ifdef __cplusplus
}
#endif

#endif /* EPIC2D_H5_H_ */
