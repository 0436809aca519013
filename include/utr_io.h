/*
 * utr_io.h -- C ABI of the host-side file formats of the UTR command-line
 * program (libutr_io.so; plain C++, no GPU).  It lets a caller prepare the
 * inputs of `utr` and read back its outputs in the formats UTR.exe uses.
 * Every function returns 0 on success and -1 on failure, with the message
 * in utr_io_last_error() (thread-local).
 */
#ifndef UTR_IO_H
#define UTR_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Reads one image; the format follows the extension: .grd ("DSBB"/"DSAA",
 * ReadmeUTR.txt:561-579), .tif/.tiff (uncompressed grey), .raw (header bytes,
 * element 4 or 8 bytes, endianness as parameter 6 of UTR.txt; nx, ny give
 * its size).  Replaces XArData::ReadFileGRD / ReadFileRAW / TIFFReadFile
 * (UTR.cpp:918-920).  cap: capacity of out in floats; the size read is
 * returned in *out_nx, *out_ny.  With out == NULL only the size is returned.
 */
int utr_io_read_image(const char* path, long nx, long ny, long header, long element,
                      int big_endian, float* out, long cap, long* out_nx, long* out_ny);

/*
 * Writes one slice; the format follows the extension: .grd (binary "DSBB",
 * with the x/y extents given), .tif/.tiff (32-bit float, or 16-bit when
 * maxout > 0 with the clamp-and-scale of ReadmeUTR.txt:448-454), .dcm (16-bit
 * DICOM; maxout must be > 0).  Replaces one slice of WriteFileStackGRD /
 * TIFFWriteFileStack / WriteFileStackDICOM (UTR.cpp:1526-1541).
 */
int utr_io_write_slice(const char* path, long nx, long ny, const float* v, double xlo, double xhi,
                       double ylo, double yhi, float minin, float maxin, long maxout);

/*
 * File name number `index` of a template with `count` entries (the
 * numbering of FileNames, ReadmeUTR.txt:241-255): written to out (cap bytes).
 */
int utr_io_file_name(const char* tmpl, long count, long index, char* out, long cap);

/*
 * Peak localization of FindPeaks.cpp:14-171 (UTR parameters 24, 25) on an
 * nz*ny*nx volume with extent6 = {xlo, xhi, ylo, yhi, zlo, zhi}: vol is
 * replaced by the isolated peaks; their (i, j, k) indices and heights, in
 * descending height, go to out_ijk (3 ints each) and out_height (up to cap
 * entries; either may be NULL); *npeaks receives their number.
 */
int utr_io_find_peaks(float* vol, long nx, long ny, long nz, const double* extent6, double atom_xy,
                      double atom_z, int* out_ijk, float* out_height, long cap, long* npeaks);

/*
 * Atom matching of pdb-compare (MultisliceCppTest/pdb-compare/
 * pdb-compare.cpp:255-433): test atoms (e.g. peaks of a reconstruction) are
 * matched to reference atoms (the specimen), both taken in descending
 * atomic number.  hungarian != 0: the one-to-one matching of least total
 * squared distance (the reference's HungarianAlgorithm::Solve, no distance
 * limit); else each reference atom in turn takes the closest unmatched test
 * atom nearer than max_dist.  match[i] / dist[i] (nref each) receive the
 * test index and distance for reference atom i, or -1; stats (6) the
 * number matched, false negatives, false positives, and the mean, maximum
 * and standard deviation (n - 1) of the matched distances.  Hydrogen in the
 * reference set is an error, as in the reference.
 */
int utr_io_match_atoms(int ntest, const double* test_xyz, const int* test_z, int nref, const double* ref_xyz,
                       const int* ref_z, double max_dist, int hungarian, int* match, double* dist,
                       double* stats);

/*
 * Host forms of the output post-processing of UTR.cpp:1383 and 1449-1498 (the
 * command-line program runs the GPU forms of include/utr_hip.h; these are
 * their host restatements, used to check them): ThresholdLow in place;
 * Reslice with flips (modes 0-5 as utr_reslice_device); thick MIP (1) or
 * average (2) slices over step input slices; 16-bit conversion.
 */
int utr_io_threshold(float* vol, long n, float background, float threshold);
int utr_io_reslice(const float* in, float* out, long nx, long ny, long nz, int mode, int flipx, int flipy,
                   int flipz);
int utr_io_thick_slices(const float* in, float* out, long nx, long ny, long nz, int mode, int step);
int utr_io_convert16(const float* in, unsigned short* out, long n, float minin, float maxin, long maxout);

const char* utr_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* UTR_IO_H */
