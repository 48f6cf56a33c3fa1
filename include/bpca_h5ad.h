/*
 * bpca_h5ad.h -- C ABI of libbrisc_h5ad.so: reading and writing the parts of
 * an .h5ad file that the PCA path uses (X, obs / var columns, obsm arrays),
 * as brisc's SingleCell loads and saves them (brisc/single_cell.py:83-1158,
 * 3954; brisc/hdf5.pyx:163 read_all_datasets reads the same datasets), on
 * the HDF5 C library.  Host only.  Every function returns 0 on success, a
 * negative value on failure (bh5_last_error() gives the message), or 1 where
 * stated (an attribute that is absent).
 *
 * Numeric datasets are read by slices of their first dimension into caller
 * buffers, in the dataset's own element type (class and size from
 * bh5_dataset_info), so that a caller can stream X through pinned buffers
 * to the device without holding it whole in host memory.
 */
#ifndef BPCA_H5AD_H
#define BPCA_H5AD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element classes. */
enum bh5_class { BH5_INT = 0, BH5_UINT = 1, BH5_FLOAT = 2, BH5_BOOL = 3, BH5_STRING = 4 };
/* Object kinds (bh5_kind). */
enum bh5_kind { BH5_MISSING = 0, BH5_GROUP = 1, BH5_DATASET = 2 };

typedef struct bh5_file bh5_file;

const char *bh5_last_error(void);
int bh5_open(const char *path, int writable, bh5_file **out);
int bh5_create(const char *path, bh5_file **out);
int bh5_close(bh5_file *f);
/* BH5_MISSING, BH5_GROUP or BH5_DATASET (negative on error). */
int bh5_kind(bh5_file *f, const char *path);
/* ndim (0..2), dims[2], element class and size in bytes (0 for
 * variable-length strings). */
int bh5_dataset_info(bh5_file *f, const char *path, int *ndim, int64_t *dims,
                     int *cls, int *size);
/* Rows [first, first + count) of a numeric or boolean dataset into out, in
 * the dataset's element type (booleans as one byte, 0 or 1). */
int bh5_read_rows(bh5_file *f, const char *path, int64_t first, int64_t count,
                  void *out);
/* A 1-D string dataset as one buffer with offsets (count + 1 entries);
 * with buffer or offsets NULL, only *total (bytes) is set. */
int bh5_read_strings(bh5_file *f, const char *path, char *buffer,
                     int64_t buffer_size, int64_t *offsets, int64_t *total);
/* String attribute (1 if absent). */
int bh5_attr_string(bh5_file *f, const char *path, const char *name, char *out,
                    int64_t out_size);
/* Integer attribute values (1 if absent). */
int bh5_attr_int64(bh5_file *f, const char *path, const char *name,
                   int64_t *out, int64_t max_count, int64_t *count);
/* Member names of a group, each followed by a newline; *needed = bytes. */
int bh5_list(bh5_file *f, const char *path, char *buffer, int64_t buffer_size,
             int64_t *needed);
/* Writers: a numeric / boolean array (replacing one at that path, creating
 * parent groups, with encoding-type / encoding-version attributes when
 * encoding_type is not NULL), a string array, a group with an encoding
 * type, and attributes. */
int bh5_write_array(bh5_file *f, const char *path, const void *data, int ndim,
                    const int64_t *dims, int cls, int size,
                    const char *encoding_type);
int bh5_write_strings(bh5_file *f, const char *path, const char *buffer,
                      const int64_t *offsets, int64_t count);
int bh5_make_group(bh5_file *f, const char *path, const char *encoding_type);
int bh5_set_attr_string(bh5_file *f, const char *path, const char *name,
                        const char *value);
int bh5_set_attr_strings(bh5_file *f, const char *path, const char *name,
                         const char *buffer, const int64_t *offsets,
                         int64_t count);
int bh5_set_attr_int64(bh5_file *f, const char *path, const char *name,
                       const int64_t *values, int64_t count);

#ifdef __cplusplus
}
#endif

#endif /* BPCA_H5AD_H */
