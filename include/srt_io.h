/*
 * srt_io.h -- C ABI of the host I/O library (libsrt_io.so) around the ray
 * kernel: the HDF5 snapshot readers and the HEALPix FITS map writer used by
 * the command-line tool.  Every function returns 0 on success and 1 on
 * failure, with a message copied into err (errlen bytes, may be null).
 */
#ifndef SRT_IO_H
#define SRT_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads a rank-3 numeric dataset (the Weyl potential) converted to float.
 * Replaces load_hdf5_array (snapshot-raytracer.cu:22-87).  dims receives the
 * three HDF5 dimensions (slowest first); out may be null to query dims only,
 * else it must hold capacity >= dims[0]*dims[1]*dims[2] floats. */
int srtio_read_scalar_field(const char* file, const char* dataset, float* out, uint64_t capacity, uint64_t* dims,
                            char* err, size_t errlen);

/* Reads a rank-3 dataset of H5T_ARRAY{3} float elements (the shift), as
 * interleaved floats.  Replaces load_hdf5_vector_field (.cu:89-186).  out may
 * be null; else capacity >= 3 * dims[0]*dims[1]*dims[2]. */
int srtio_read_vector_field(const char* file, const char* dataset, float* out, uint64_t capacity, uint64_t* dims,
                            char* err, size_t errlen);

/* Writes a full-sky HEALPix map (12 nside^2 floats) as a FITS binary table with
 * the HEALPix keywords.  Replaces chealpix write_healpix_map(map, nside, file,
 * nest, coordsys) as called at .cu:947-987 (nest = 0, coordsys "C"). */
int srtio_write_healpix_map(const char* file, const float* map, int64_t nside, int nested, const char* coordsys,
                            char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* SRT_IO_H */
