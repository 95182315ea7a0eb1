/*
 * HDF5 files of a simulation in the OPS layout (include/opensbli_amd_h5.h).
 *
 * Host code over the HDF5 C library.  Every call opens the file, does its
 * work and closes every handle it opened, so a file is complete on disk when
 * a call returns.
 */
#include "../../include/opensbli_amd_h5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static __thread char g_err[512];

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char* osbl_h5_last_error(void) { return g_err; }

/* the library prints its error stack by default; failures are reported through the return codes instead */
static void quiet(void) {
  static int done = 0;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    done = 1;
  }
}

int osbl_h5_library_version(void) {
  unsigned maj = 0, min = 0, rel = 0;
  if (H5get_libversion(&maj, &min, &rel) < 0) return fail(OSBL_H5_E_IO, "H5get_libversion failed");
  return (int)(maj * 10000 + min * 100 + rel);
}

/* handles opened by one call, closed in reverse order by done() */
typedef struct {
  hid_t file, group, dset, space, attr, type;
} Handles;

static void done(Handles* h) {
  if (h->type >= 0) H5Tclose(h->type);
  if (h->attr >= 0) H5Aclose(h->attr);
  if (h->space >= 0) H5Sclose(h->space);
  if (h->dset >= 0) H5Dclose(h->dset);
  if (h->group >= 0) H5Gclose(h->group);
  if (h->file >= 0) H5Fclose(h->file);
}

#define HANDLES_INIT {-1, -1, -1, -1, -1, -1}

static int write_int_attr(hid_t obj, const char* name, const int32_t* v, int n) {
  if (H5Aexists(obj, name) > 0 && H5Adelete(obj, name) < 0) return fail(OSBL_H5_E_IO, "cannot replace attribute %s", name);
  hsize_t dims[1] = {(hsize_t)n};
  hid_t sp = H5Screate_simple(1, dims, NULL);
  if (sp < 0) return fail(OSBL_H5_E_IO, "H5Screate_simple failed");
  hid_t a = H5Acreate2(obj, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT);
  int rc = OSBL_H5_OK;
  if (a < 0 || H5Awrite(a, H5T_NATIVE_INT32, v) < 0) rc = fail(OSBL_H5_E_IO, "cannot write attribute %s", name);
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return rc;
}

/* a scalar fixed-length string of the given width, null padded */
static int write_str_attr(hid_t obj, const char* name, const char* s, size_t width) {
  if (strlen(s) > width) return fail(OSBL_H5_E_ARG, "attribute %s: '%s' is longer than %zu characters", name, s, width);
  if (H5Aexists(obj, name) > 0 && H5Adelete(obj, name) < 0) return fail(OSBL_H5_E_IO, "cannot replace attribute %s", name);
  char buf[64];
  memset(buf, 0, sizeof buf);
  memcpy(buf, s, strlen(s));
  hid_t t = H5Tcopy(H5T_C_S1);
  hid_t sp = H5Screate(H5S_SCALAR);
  int rc = OSBL_H5_OK;
  if (t < 0 || sp < 0 || H5Tset_size(t, width) < 0 || H5Tset_strpad(t, H5T_STR_NULLPAD) < 0) {
    rc = fail(OSBL_H5_E_IO, "cannot make the string type of attribute %s", name);
  } else {
    hid_t a = H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    if (a < 0 || H5Awrite(a, t, buf) < 0) rc = fail(OSBL_H5_E_IO, "cannot write attribute %s", name);
    if (a >= 0) H5Aclose(a);
  }
  if (sp >= 0) H5Sclose(sp);
  if (t >= 0) H5Tclose(t);
  return rc;
}

/* an integer attribute of exactly n entries */
static int read_int_attr(hid_t obj, const char* obj_name, const char* name, int32_t* out, int n) {
  if (H5Aexists(obj, name) <= 0) return fail(OSBL_H5_E_FORMAT, "%s has no attribute %s", obj_name, name);
  hid_t a = H5Aopen(obj, name, H5P_DEFAULT);
  if (a < 0) return fail(OSBL_H5_E_IO, "cannot open attribute %s of %s", name, obj_name);
  hid_t sp = H5Aget_space(a);
  hssize_t np = sp >= 0 ? H5Sget_simple_extent_npoints(sp) : -1;
  int rc = OSBL_H5_OK;
  if (np != n) {
    rc = fail(OSBL_H5_E_FORMAT, "attribute %s of %s has %lld entries, expected %d", name, obj_name, (long long)np, n);
  } else {
    hid_t t = H5Aget_type(a);
    if (t < 0 || H5Tget_class(t) != H5T_INTEGER) rc = fail(OSBL_H5_E_FORMAT, "attribute %s of %s is not integer", name, obj_name);
    else if (H5Aread(a, H5T_NATIVE_INT32, out) < 0) rc = fail(OSBL_H5_E_IO, "cannot read attribute %s of %s", name, obj_name);
    if (t >= 0) H5Tclose(t);
  }
  if (sp >= 0) H5Sclose(sp);
  H5Aclose(a);
  return rc;
}

static int check_name(const char* what, const char* s) {
  if (!s || !*s) return fail(OSBL_H5_E_ARG, "%s name is empty", what);
  if (strchr(s, '/')) return fail(OSBL_H5_E_ARG, "%s name '%s' contains '/'", what, s);
  return OSBL_H5_OK;
}

static int open_group(Handles* h, const char* path, const char* block_name, unsigned flags) {
  if (access(path, F_OK) != 0) return fail(OSBL_H5_E_MISSING, "%s: no such file", path);
  h->file = H5Fopen(path, flags, H5P_DEFAULT);
  if (h->file < 0) return fail(OSBL_H5_E_IO, "%s: not an HDF5 file or cannot be opened", path);
  if (H5Lexists(h->file, block_name, H5P_DEFAULT) <= 0)
    return fail(OSBL_H5_E_MISSING, "%s: no block group '%s'", path, block_name);
  h->group = H5Gopen2(h->file, block_name, H5P_DEFAULT);
  if (h->group < 0) return fail(OSBL_H5_E_FORMAT, "%s: '%s' is not a group", path, block_name);
  return OSBL_H5_OK;
}

int osbl_h5_fetch_block(const char* path, const char* block_name, int32_t block_index, int32_t ndim,
                        int32_t truncate) {
  quiet();
  if (!path) return fail(OSBL_H5_E_ARG, "null path");
  if (check_name("block", block_name)) return OSBL_H5_E_ARG;
  if (ndim < 1 || ndim > 3) return fail(OSBL_H5_E_ARG, "ndim %d: 1, 2 or 3", ndim);
  if (block_index < 0) return fail(OSBL_H5_E_ARG, "block index %d < 0", block_index);
  Handles h = HANDLES_INIT;
  int rc = OSBL_H5_OK;
  if (truncate || access(path, F_OK) != 0) h.file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  else h.file = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (h.file < 0) {
    rc = fail(OSBL_H5_E_IO, "%s: cannot create or open for writing", path);
    goto out;
  }
  if (H5Lexists(h.file, block_name, H5P_DEFAULT) > 0) h.group = H5Gopen2(h.file, block_name, H5P_DEFAULT);
  else h.group = H5Gcreate2(h.file, block_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (h.group < 0) {
    rc = fail(OSBL_H5_E_IO, "%s: cannot create group '%s'", path, block_name);
    goto out;
  }
  if ((rc = write_int_attr(h.group, "dims", &ndim, 1))) goto out;
  if ((rc = write_str_attr(h.group, "ops_type", "ops_block", 9))) goto out;
  if ((rc = write_int_attr(h.group, "index", &block_index, 1))) goto out;
out:
  done(&h);
  return rc;
}

int osbl_h5_fetch_dat(const char* path, const char* block_name, int32_t block_index, const char* dat_name,
                      int32_t ndim, const int32_t* size, const int32_t* d_m, const int32_t* d_p,
                      const double* data) {
  quiet();
  if (!path || !size || !d_m || !d_p || !data) return fail(OSBL_H5_E_ARG, "null argument");
  if (check_name("block", block_name) || check_name("dataset", dat_name)) return OSBL_H5_E_ARG;
  if (ndim < 1 || ndim > 3) return fail(OSBL_H5_E_ARG, "ndim %d: 1, 2 or 3", ndim);
  hsize_t dims[3];
  for (int d = 0; d < ndim; ++d) {
    if (size[d] < 1) return fail(OSBL_H5_E_ARG, "size[%d] = %d < 1", d, size[d]);
    if (d_m[d] > 0 || d_p[d] < 0) return fail(OSBL_H5_E_ARG, "halo depths of direction %d: d_m %d must be <= 0, d_p %d >= 0", d, d_m[d], d_p[d]);
    dims[ndim - 1 - d] = (hsize_t)(size[d] - d_m[d] + d_p[d]);
  }
  Handles h = HANDLES_INIT;
  int rc = open_group(&h, path, block_name, H5F_ACC_RDWR);
  if (rc) {
    if (rc == OSBL_H5_E_MISSING) fail(rc, "%s (osbl_h5_fetch_block writes the block group first)", g_err);
    goto out;
  }
  if (H5Lexists(h.group, dat_name, H5P_DEFAULT) > 0 && H5Ldelete(h.group, dat_name, H5P_DEFAULT) < 0) {
    rc = fail(OSBL_H5_E_IO, "%s: cannot replace dataset %s", path, dat_name);
    goto out;
  }
  h.space = H5Screate_simple(ndim, dims, NULL);
  h.dset = h.space < 0 ? -1 : H5Dcreate2(h.group, dat_name, H5T_IEEE_F64LE, h.space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (h.dset < 0 || H5Dwrite(h.dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
    rc = fail(OSBL_H5_E_IO, "%s: cannot write dataset %s", path, dat_name);
    goto out;
  }
  {
    const int32_t one = 1, zeros[3] = {0, 0, 0};
    if ((rc = write_int_attr(h.dset, "d_p", d_p, ndim))) goto out;
    if ((rc = write_int_attr(h.dset, "d_m", d_m, ndim))) goto out;
    if ((rc = write_int_attr(h.dset, "dim", &one, 1))) goto out;
    if ((rc = write_str_attr(h.dset, "ops_type", "ops_dat", 10))) goto out;
    if ((rc = write_int_attr(h.dset, "block_index", &block_index, 1))) goto out;
    if ((rc = write_int_attr(h.dset, "base", zeros, ndim))) goto out;
    if ((rc = write_str_attr(h.dset, "type", "double", 15))) goto out;
    if ((rc = write_str_attr(h.dset, "block", block_name, 25))) goto out;
    if ((rc = write_int_attr(h.dset, "size", size, ndim))) goto out;
  }
out:
  done(&h);
  return rc;
}

static int open_dat(Handles* h, const char* path, const char* block_name, const char* dat_name) {
  if (!path) return fail(OSBL_H5_E_ARG, "null path");
  if (check_name("block", block_name) || check_name("dataset", dat_name)) return OSBL_H5_E_ARG;
  int rc = open_group(h, path, block_name, H5F_ACC_RDONLY);
  if (rc) return rc;
  if (H5Lexists(h->group, dat_name, H5P_DEFAULT) <= 0)
    return fail(OSBL_H5_E_MISSING, "%s: block '%s' has no dataset '%s'", path, block_name, dat_name);
  h->dset = H5Dopen2(h->group, dat_name, H5P_DEFAULT);
  if (h->dset < 0) return fail(OSBL_H5_E_FORMAT, "%s: '%s' is not a dataset", path, dat_name);
  h->type = H5Dget_type(h->dset);
  if (h->type < 0 || H5Tget_class(h->type) != H5T_FLOAT)
    return fail(OSBL_H5_E_FORMAT, "%s: dataset %s is not floating point", path, dat_name);
  h->space = H5Dget_space(h->dset);
  if (h->space < 0) return fail(OSBL_H5_E_IO, "%s: cannot read the shape of %s", path, dat_name);
  return OSBL_H5_OK;
}

int osbl_h5_dat_info(const char* path, const char* block_name, const char* dat_name, int32_t* ndim,
                     int32_t* size, int32_t* d_m, int32_t* d_p) {
  quiet();
  if (!ndim || !size || !d_m || !d_p) return fail(OSBL_H5_E_ARG, "null argument");
  Handles h = HANDLES_INIT;
  int rc = open_dat(&h, path, block_name, dat_name);
  if (rc) goto out;
  {
    const int nd = H5Sget_simple_extent_ndims(h.space);
    hsize_t dims[H5S_MAX_RANK];
    if (nd < 1 || nd > 3 || H5Sget_simple_extent_dims(h.space, dims, NULL) < 0) {
      rc = fail(OSBL_H5_E_FORMAT, "%s: dataset %s has rank %d (1, 2 or 3 expected)", path, dat_name, nd);
      goto out;
    }
    int32_t s[3], m[3], p[3];
    if ((rc = read_int_attr(h.dset, dat_name, "size", s, nd))) goto out;
    if ((rc = read_int_attr(h.dset, dat_name, "d_m", m, nd))) goto out;
    if ((rc = read_int_attr(h.dset, dat_name, "d_p", p, nd))) goto out;
    for (int d = 0; d < nd; ++d) {
      /* the low-side depth is stored negative (OPS), read by magnitude as the reference's readers do
         (Sod_shock_tube/plot.py:19-23) */
      const int32_t lo = m[d] < 0 ? -m[d] : m[d];
      const int32_t hi = p[d] < 0 ? -p[d] : p[d];
      if ((hsize_t)(s[d] + lo + hi) != dims[nd - 1 - d]) {
        rc = fail(OSBL_H5_E_FORMAT,
                  "%s: dataset %s direction %d: size %d with halos %d/%d does not match its extent %llu", path,
                  dat_name, d, s[d], m[d], p[d], (unsigned long long)dims[nd - 1 - d]);
        goto out;
      }
      size[d] = s[d];
      d_m[d] = -lo;
      d_p[d] = hi;
    }
    *ndim = nd;
  }
out:
  done(&h);
  return rc;
}

int osbl_h5_read_dat(const char* path, const char* block_name, const char* dat_name, double* data,
                     int64_t count) {
  quiet();
  if (!data) return fail(OSBL_H5_E_ARG, "null data");
  Handles h = HANDLES_INIT;
  int rc = open_dat(&h, path, block_name, dat_name);
  if (rc) goto out;
  {
    const hssize_t np = H5Sget_simple_extent_npoints(h.space);
    if (np != count) {
      rc = fail(OSBL_H5_E_ARG, "%s: dataset %s holds %lld values, the buffer %lld", path, dat_name, (long long)np,
                (long long)count);
      goto out;
    }
    if (H5Dread(h.dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
      rc = fail(OSBL_H5_E_IO, "%s: cannot read dataset %s", path, dat_name);
  }
out:
  done(&h);
  return rc;
}

typedef struct {
  char* buf;
  int64_t cap, used;
  int count, overflow;
} Names;

static herr_t collect(hid_t g, const char* name, const H5L_info_t* info, void* op) {
  (void)g;
  (void)info;
  Names* n = (Names*)op;
  const int64_t len = (int64_t)strlen(name);
  if (n->used + len + 2 > n->cap) {
    n->overflow = 1;
    return 1;
  }
  memcpy(n->buf + n->used, name, (size_t)len);
  n->used += len;
  n->buf[n->used++] = '\n';
  n->buf[n->used] = 0;
  n->count++;
  return 0;
}

int osbl_h5_list(const char* path, const char* block_name, char* buf, int64_t cap) {
  quiet();
  if (!path || !buf || cap < 1) return fail(OSBL_H5_E_ARG, "null argument");
  buf[0] = 0;
  Handles h = HANDLES_INIT;
  int rc;
  hid_t where;
  if (block_name) {
    if ((rc = check_name("block", block_name))) return rc;
    if ((rc = open_group(&h, path, block_name, H5F_ACC_RDONLY))) goto out;
    where = h.group;
  } else {
    if (access(path, F_OK) != 0) {
      rc = fail(OSBL_H5_E_MISSING, "%s: no such file", path);
      goto out;
    }
    h.file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (h.file < 0) {
      rc = fail(OSBL_H5_E_IO, "%s: not an HDF5 file or cannot be opened", path);
      goto out;
    }
    where = h.file;
  }
  {
    Names n = {buf, cap, 0, 0, 0};
    hsize_t idx = 0;
    if (H5Literate(where, H5_INDEX_NAME, H5_ITER_INC, &idx, collect, &n) < 0) {
      rc = fail(OSBL_H5_E_IO, "%s: cannot list names", path);
      goto out;
    }
    rc = n.overflow ? fail(OSBL_H5_E_ARG, "%s: the names need more than %lld bytes", path, (long long)cap) : n.count;
  }
out:
  done(&h);
  return rc;
}
