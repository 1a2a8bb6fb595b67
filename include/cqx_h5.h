/* cqx_h5: ChronusQ binary-file (HDF5) interop, a host-only library of its
 * own (libcqx_h5.so, linked against the HDF5 C library) so that the J/K
 * engine itself never depends on HDF5.
 *
 * It reads and writes the files ChronusQ's SafeFile produces
 * (include/util/files.hpp:59-342): datasets addressed by path ("/SCF/FOCK_SCALAR"),
 * real data as IEEE F64LE, complex data as the compound {double "r";
 * double "i"} of H5PredType<dcomplex> (files.hpp:44-56), indices as U64LE,
 * matrices stored as ChronusQ's raw column-major buffer with dims {N, N}
 * (so the HDF5 row-major view is the transpose).  All entry points return
 * 0 on success; cqx_h5_last_error() holds the message of the last failure
 * on the calling thread. */
#ifndef CQX_H5_H
#define CQX_H5_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cqx_h5 cqx_h5;

/* open modes: SafeFile's H5F_ACC_RDONLY / H5F_ACC_RDWR / createFile()
 * (H5F_ACC_TRUNC, files.hpp:61-70) */
enum { CQX_H5_READ = 0, CQX_H5_RDWR = 1, CQX_H5_CREATE = 2 };
/* element types: F64, complex<double> as {r, i}, U64 (size_t, FIELD_TYPE) */
enum { CQX_H5_F64 = 0, CQX_H5_C128 = 1, CQX_H5_U64 = 2 };
/* status codes beyond 0 */
enum { CQX_H5_ERR_ARG = 1, CQX_H5_ERR_IO = 2, CQX_H5_ERR_MISSING = 3, CQX_H5_ERR_TYPE = 4 };

int cqx_h5_open(const char *path, int mode, cqx_h5 **out);
int cqx_h5_close(cqx_h5 *f);

/* SafeFile::getDims (files.hpp:307-322): rank and extents (at most 8) of a
 * dataset; CQX_H5_ERR_MISSING when it does not exist */
int cqx_h5_dims(cqx_h5 *f, const char *dataset, int *rank, unsigned long long *dims);
/* element type of a dataset (CQX_H5_F64 / C128 / U64) */
int cqx_h5_type(cqx_h5 *f, const char *dataset, int *dtype);

/* SafeFile::readData (files.hpp:154-158): the whole dataset into buf,
 * which holds count elements of dtype (count must equal the dataset's
 * element count; dtype must match the stored type class) */
int cqx_h5_read(cqx_h5 *f, const char *dataset, int dtype, void *buf, size_t count);

/* SafeFile::safeWriteData (files.hpp:244-268): write into the dataset if it
 * exists (its extents must equal dims), else create it, creating the
 * missing groups of its path first */
int cqx_h5_write(cqx_h5 *f, const char *dataset, int dtype, int rank, const unsigned long long *dims,
                 const void *buf);

const char *cqx_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
