// utr_io.hpp -- host-side file formats and parameter grammar of the UTR
// command-line program (the parts of UTR.cpp:55-784 and 1281-1547 that do
// not run on the GPU).  Plain C++17, no HIP.
//
// Reference: MultisliceCppTest/PhaseRetrieval/UTR.cpp, ReadmeUTR.txt.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace utrio {

// One row of an orientation file: Z and Y' rotation (degrees) and the
// (Z'' rotation, defocus) pairs that follow (ReadmeUTR.txt:186-198).
struct Orientation {
    double angZ = 0, angY = 0;
    std::vector<std::pair<double, double>> defocus;  // (Z'' degrees, defocus UL)
    double shift_x = 0, shift_y = 0;                 // .relionnew only
    double astig_dz = 0, astig_angle = 0;            // .relionnew only
};

// All parameters of a UTR.txt (35 lines) or .xri (27 lines) file, with the
// derived geometry of UTR.cpp:284-388.
struct Config {
    bool xri = false;
    int modality = 0;                      // 1
    bool verbose = true;                   // 2
    std::string defocus_file;              // 3
    std::string subseries_file = "NONE";   // 4
    std::string input_template;            // 5
    long raw_header = 0;                   // 6
    bool raw_big_endian = false;
    long raw_element = 4;
    double f1 = 1.0, f2 = 0.0;             // 7
    long nx = 0, ny = 0;                   // 8
    double xst = 0;                        // 9
    double zlo = 0, zhi = 0, dzextra = 0;  // 10
    double dxc = 0, dyc = 0, dzc = 0;      // 11
    double wl = 0;                         // 12
    double aobj = 0;                       // 13 (mrad)
    double cs3_mm = 0, cs5_mm = 0;         // 14
    int ctf_mode = 0;                      // 15
    int escc = 0;                          // 16
    double asigma = 0;                     // 17 (beta/delta; -1/(delta/beta) for .xri)
    double nsr = 0;                        // 18
    double epsilon = 0;                    // 19
    int symmetry_mode = 0;                 // 20
    std::string symmetry_file;             // 21
    int inv_laplace = 0;                   // 22
    double lp_filter = 0, background = 0, threshold = 0;  // 23
    int peaks = 0;                         // 24
    double atom_xy = 0, atom_z = 0;        // 25
    std::string output_template;           // 26
    float b16_minin = 0, b16_maxin = 0;    // 27
    long b16_maxout = 0;
    double rot_out[3] = {0, 0, 0};         // 28 (as written; the program negates them)
    int reslice = 0;                       // 29
    bool flip[3] = {false, false, false};  // 30
    int thick_mode = 0;                    // 31
    double thick_UL = 0;
    int save_sampling = 0;                 // 32
    int reprocess = 0;                     // 33
    std::string aux_folder;                // 34
    int threads = 1;                       // 35
    // derived (UTR.cpp:290-321)
    long nz = 0;
    double xlo = 0, xhi = 0, ylo = 0, yhi = 0;
    bool out16() const { return !(b16_minin == 0 && b16_maxin == 0 && b16_maxout == 0); }
};

// Upper-case extension with the dot (".GRD"), "" if none.
std::string file_extension(const std::string& path);
// Folder part of a path ("." if none).
std::string path_folder(const std::string& path);

// Reads and validates a parameter file; throws std::runtime_error with the
// reference's messages.  Folder existence checks are left to the caller.
// tables (may be empty) runs right after parameter 4 is read, where
// UTR.cpp:113-230 reads the orientation and subseries tables and validates
// them, so that their errors come before those of parameters 5 and later.
Config read_config(const std::string& path, const std::function<void(const Config&)>& tables = {});

// Orientation files: ".txt" (read as written) and ".relionnew" (all values
// except the astigmatism angle negated, ReadmeUTR.txt:215-220).
std::vector<Orientation> read_orientations(const std::string& path);

// Subseries selection file: "index weight" per line.
void read_subseries(const std::string& path, std::vector<int>& idx, std::vector<float>& w);

// Output / input file names from a template (ReadmeUTR.txt:241-255, 429-436):
// the index is inserted before the extension.  If the template's name ends
// in digits, numbering starts at that number with at least that many digits;
// otherwise it starts at 0, zero-padded to the digits of count-1.
std::vector<std::string> file_names(const std::string& tmpl, long count);

// Images: a 2D float array, row-major (x fastest).
struct Image {
    long nx = 0, ny = 0;
    std::vector<float> v;
    double xlo = 0, xhi = 0, ylo = 0, yhi = 0;  // GRD header
};

Image read_grd(const std::string& path);          // "DSBB" binary or "DSAA" text
void write_grd(const std::string& path, long nx, long ny, const float* v, double xlo,
               double xhi, double ylo, double yhi);
Image read_raw(const std::string& path, long nx, long ny, long header, long element,
               bool big_endian);
Image read_tiff(const std::string& path);         // uncompressed grey TIFF (8/16/32-bit int, 32/64-bit float)
void write_tiff_f32(const std::string& path, long nx, long ny, const float* v);
void write_tiff_u16(const std::string& path, long nx, long ny, const uint16_t* v);
void write_dicom_u16(const std::string& path, long nx, long ny, const uint16_t* v,
                     const std::map<std::string, std::string>& tags, long instance,
                     double pixel_spacing_mm);

// Default DICOM tags (UTR.cpp:600-611) and the UTR_DICOM.txt reader.
std::map<std::string, std::string> dicom_tags(const std::string& path, bool* found);

// Centred pad (value) or trim to (ny, nx) as UTR.cpp:928-947; returns
// +1 padded, -1 trimmed, 0 unchanged; throws if one dimension grows and the
// other shrinks.
int pad_or_trim(Image& im, long nx, long ny, float pad_value);

// 16-bit conversion (ReadmeUTR.txt:448-454): clamp to [minin, maxin], map
// linearly onto [0, maxout], round to nearest.
void to_u16(const float* v, size_t n, float minin, float maxin, long maxout, uint16_t* out);

// Volume post-processing on the host (UTR.cpp:1383-1490).
struct Volume {
    long nx = 0, ny = 0, nz = 0;
    std::vector<float> v;
    float& at(long k, long j, long i) { return v[((size_t)k * ny + j) * nx + i]; }
    float at(long k, long j, long i) const { return v[((size_t)k * ny + j) * nx + i]; }
};
void threshold_low(Volume& vol, float a, float b);
Volume reslice(const Volume& vol, int mode, bool fx, bool fy, bool fz);
Volume thick_slices(const Volume& vol, int mode, int step);

// One located peak: grid indices, coordinates and height.
struct Peak {
    int i, j, k;
    float x, y, z, height;
};
// Peak localization of FindPeaks.cpp:14-171 (parameters 24, 25): at most one
// peak per atom-sized box inside the inscribed sphere, the smaller of any two
// peaks closer than atom_xy removed; vol is replaced by the isolated peaks.
std::vector<Peak> find_peaks(Volume& vol, double xlo, double xhi, double ylo, double yhi, double zlo,
                             double zhi, double atom_xy, double atom_z);
// Kirkland's XYZ format: title, cell a b c, "Z x y z occupancy wobble" per
// atom (carbon, occupancy = peak height, wobble 0), "-1".
void write_xyz(const std::string& path, float a, float c, const std::vector<Peak>& peaks);

}  // namespace utrio
