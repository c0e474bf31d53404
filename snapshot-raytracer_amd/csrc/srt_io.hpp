// Host I/O either side of the ray kernel: HDF5 snapshot readers and the
// HEALPix FITS map writer.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace srt_io {

// Where a reader puts the floats: called once the dimensions are known with
// the number of floats the dataset holds; returns the buffer to fill, or null
// to read the dimensions only.
using FloatSink = std::function<float*(uint64_t count)>;

// Scalar dataset (the Weyl potential): any numeric type, converted to float
// by HDF5, rank 3 required.  dims in HDF5 order (slowest first).
// Mirrors load_hdf5_array (snapshot-raytracer.cu:22-87); the reference does
// not check the rank.  Returns false and fills err on failure.
//
// Both readers take the floats straight from the file when the dataset stores
// them as they are returned (contiguous layout, no filters, no user block,
// little-endian IEEE binary32 elements, or arrays of 3 of them): the bytes are
// read with pread from several threads at the dataset's file offset.  Any
// other dataset is read and converted by H5Dread.
bool read_scalar_field(const std::string& file, const std::string& dataset, const FloatSink& sink,
                       std::vector<uint64_t>& dims, std::string& err);

// Vector dataset (the shift): rank-3 dataspace whose element type is a
// one-dimensional H5T_ARRAY of length 3, read as float[3] per voxel,
// interleaved.  Mirrors load_hdf5_vector_field (.cu:89-186).
bool read_vector_field(const std::string& file, const std::string& dataset, const FloatSink& sink,
                       std::vector<uint64_t>& dims, std::string& err);

// Full-sky HEALPix map as a FITS binary table: primary HDU without data, then a
// BINTABLE with one float32 column of 12 nside^2 rows and the HEALPix keywords
// (PIXTYPE, ORDERING, NSIDE, FIRSTPIX, LASTPIX, INDXSCHM, OBJECT, COORDSYS).
// Stands in for chealpix write_healpix_map(map, nside, name, nest, coordsys)
// as called at .cu:947 (nest = 0, coordsys "C").  Overwrites the file.
bool write_healpix_map(const std::string& file, const float* map, int64_t nside, bool nested,
                       const std::string& coordsys, std::string& err);

}  // namespace srt_io
