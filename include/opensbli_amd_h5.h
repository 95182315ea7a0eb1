/*
 * opensbli_amd_h5.h -- C-ABI of libosbl_h5.so: the HDF5 files of an OpenSBLI
 * simulation, in the layout the OPS library reads and writes.
 *
 * A file holds one group per block (named by the block, "opensbliblock00"),
 * with the attributes ops_type = "ops_block", dims = [ndim] and
 * index = [block number]; in it one float64 dataset per array (named
 * "<array>_B<block number>", e.g. "rho_B0") holding the array WITH its halo
 * layers, dimensions in reverse order (x varies fastest), and the attributes
 * d_p, d_m, dim, ops_type = "ops_dat", block_index, base, type = "double",
 * block and size -- the set the reference's own writer gives a dataset
 * (helperfunctions.py:80-101), with size the interior points per direction
 * (helperfunctions.py:99 with kelvin_helmholtz.py:100-105) and d_m <= 0 the
 * low-side halo depths.  Fixed-length string attributes use the widths of
 * that writer (S9, S10, S15, S25, null padded).
 *
 * The generated program calls the OPS runtime for this; the calls replaced:
 *   osbl_h5_fetch_block  <- ops_fetch_block_hdf5_file(block, file)  (io_hdf5.py:126)
 *   osbl_h5_fetch_dat    <- ops_fetch_dat_hdf5_file(dat, file)      (io_hdf5.py:121-123)
 *   osbl_h5_dat_info,
 *   osbl_h5_read_dat     <- ops_decl_dat_hdf5(block, 1, type, name, file)  (opsc.py:702-705)
 *   osbl_h5_list         <- the dataset walk of apps/check.py:7-17
 *
 * Conventions as in opensbli_amd.h: 0 on success, a negative OSBL_H5_E_* code
 * on failure with a thread-local message from osbl_h5_last_error(); buffers
 * are borrowed for the call.  Host code only: no device memory is touched.
 */
#ifndef OPENSBLI_AMD_H5_H
#define OPENSBLI_AMD_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSBL_H5_OK 0
#define OSBL_H5_E_ARG (-1)      /* invalid argument */
#define OSBL_H5_E_IO (-2)       /* the HDF5 library failed (open, create, read, write) */
#define OSBL_H5_E_MISSING (-3)  /* file, block group or dataset not found */
#define OSBL_H5_E_FORMAT (-4)   /* present but not in the OPS layout (attributes, shape, type) */

/* Thread-local message describing the last failure. */
const char* osbl_h5_last_error(void);

/* Version of the HDF5 library in use, as major*10000 + minor*100 + release. */
int osbl_h5_library_version(void);

/*
 * Create the block's group with its attributes.  truncate != 0 creates the
 * file anew; otherwise an existing file is opened and kept.  An existing group
 * of that name has its attributes rewritten.
 */
int osbl_h5_fetch_block(const char* path, const char* block_name, int32_t block_index, int32_t ndim,
                        int32_t truncate);

/*
 * Write one array into the block's group (created by osbl_h5_fetch_block
 * first), replacing a dataset of the same name.  size[d] are the interior
 * points, d_m[d] <= 0 and d_p[d] >= 0 the halo depths, per direction d
 * (x first); data holds prod(size[d] - d_m[d] + d_p[d]) doubles, x fastest.
 */
int osbl_h5_fetch_dat(const char* path, const char* block_name, int32_t block_index, const char* dat_name,
                      int32_t ndim, const int32_t* size, const int32_t* d_m, const int32_t* d_p,
                      const double* data);

/*
 * Shape of one dataset: ndim and, per direction (x first), the interior
 * points and halo depths from its attributes, checked against the dataset's
 * dimensions.  size, d_m and d_p have room for 3 entries.
 */
int osbl_h5_dat_info(const char* path, const char* block_name, const char* dat_name, int32_t* ndim,
                     int32_t* size, int32_t* d_m, int32_t* d_p);

/* Read one dataset, halos included, into data (count doubles, which must equal
   the dataset's size). */
int osbl_h5_read_dat(const char* path, const char* block_name, const char* dat_name, double* data,
                     int64_t count);

/*
 * Names of the groups of the file (block_name NULL) or of the datasets of a
 * group, each followed by '\n', into buf (cap bytes, NUL-terminated).  Returns
 * the number of names, or OSBL_H5_E_ARG when buf is too small.
 */
int osbl_h5_list(const char* path, const char* block_name, char* buf, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif
