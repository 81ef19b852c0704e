/*
 * ptycxi -- the reference's CXI-style result files through the HDF5 C library.
 *
 * The reference writes and reads its results with h5py
 * (ptychography/display.py:21-56 write_cxi, :59-117 read_cxi): groups
 * "input" (I, M, P, O, R, B) and "output" (I, P, O, R, B, eMod), I
 * gzip-compressed.  These entry points write and read one dataset at a time,
 * in the layout h5py gives the same numpy arrays, so files written here open
 * in the reference's read_cxi and files it wrote open here:
 *   - complex arrays are compound types {"r", "i"} of float32 / float64;
 *   - bool arrays are an int8 enum {FALSE = 0, TRUE = 1};
 *   - other types are little-endian IEEE floats and signed integers;
 *   - a scalar (the reference's mask 1) is a dataset of rank 0.
 * Every function opens and closes the file itself and returns PTY_OK (0) or a
 * negative code (PTY_EINVAL, PTY_EHIP = an HDF5 call failed, PTY_ENOENT = no
 * such dataset); pty_cxi_last_error() describes the last failure on the
 * calling thread.  No GPU is involved.
 */
#ifndef PTYCXI_H
#define PTYCXI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PTY_CXI_F32 = 1,
  PTY_CXI_F64 = 2,
  PTY_CXI_C64 = 3,    /* complex64: {r, i} float32 */
  PTY_CXI_C128 = 4,   /* complex128: {r, i} float64 */
  PTY_CXI_I32 = 5,
  PTY_CXI_I64 = 6,
  PTY_CXI_BOOL = 7    /* one byte per value, 0 or 1 */
};
enum { PTY_CXI_ENOENT = -6, PTY_CXI_MAXDIM = 8 };

const char* pty_cxi_last_error(void);

/* Create (or truncate) the file at path: write_cxi replaces an existing file
 * (display.py:13-19). */
int pty_cxi_create(const char* path);

/* Write dataset `name` ("input/I", "output/eMod", ...; missing groups are
 * created) of ndim (0..8) dims from host memory in layout dtype.
 * gzip_level 1..9 stores it chunked with the deflate filter (one chunk per
 * few frames along the first axis); 0 stores it contiguous. */
int pty_cxi_write(const char* path, const char* name, int32_t dtype, int32_t ndim,
                  const int64_t* dims, const void* data, int32_t gzip_level);

/* Type and shape of dataset `name`: *dtype one of PTY_CXI_*, dims[0..ndim).
 * PTY_CXI_ENOENT if the file has no such dataset. */
int pty_cxi_info(const char* path, const char* name, int32_t* dtype, int32_t* ndim, int64_t* dims);

/* Read dataset `name` into host memory as dtype (the library converts between
 * float widths and integer widths); max_rows >= 0 reads only the first
 * min(max_rows, dims[0]) entries along the first axis (read_cxi's maxlen,
 * display.py:68,87), -1 reads everything. */
int pty_cxi_read(const char* path, const char* name, int32_t dtype, void* data, int64_t max_rows);

#ifdef __cplusplus
}
#endif
#endif /* PTYCXI_H */
