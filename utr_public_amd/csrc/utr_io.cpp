// utr_io.cpp -- parameter grammar, orientation tables, image formats and
// host post-processing of the UTR command-line program.
//
// Reference: MultisliceCppTest/PhaseRetrieval/UTR.cpp:55-784 (parameter
// file), ReadmeUTR.txt:186-255 (orientation and subseries files, file-name
// templates), ReadmeUTR.txt:561-579 (GRD), UTR.cpp:1383-1547 (threshold,
// reslice, thick slices, 16-bit output, writers).
#include "utr_io.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace utrio {

namespace {

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> t;
    std::istringstream is(s);
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

// Line-oriented reader of a parameter file: skips the leading "//" comment
// lines (UTR.cpp:87-92), then hands out one line per parameter.
struct ParamReader {
    std::ifstream f;
    std::string first;
    bool have_first = false;
    explicit ParamReader(const std::string& path) : f(path) {
        if (!f) fail("Error: cannot open parameter file " + path + ".");
        std::string line;
        while (std::getline(f, line)) {
            if (!(line.size() >= 2 && line[0] == '/' && line[1] == '/')) {
                first = line;
                have_first = true;
                return;
            }
        }
    }
    // tokens of the next parameter line; at least `need` of them (title included)
    std::vector<std::string> next(size_t need, const std::string& err) {
        std::string line;
        if (have_first) {
            line = first;
            have_first = false;
        } else if (!std::getline(f, line)) {
            fail(err);
        }
        auto t = split_ws(line);
        if (t.size() < need) fail(err);
        return t;
    }
};

double num(const std::string& s) { return std::atof(s.c_str()); }
long inum(const std::string& s) { return std::atol(s.c_str()); }

bool starts_comment(const std::string& l) {
    size_t p = l.find_first_not_of(" \t\r");
    return p == std::string::npos || (l.compare(p, 2, "//") == 0);
}

uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }
uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
uint64_t bswap64(uint64_t v) {
    return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}

std::vector<char> read_all(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("Error: cannot open input file " + path + ".");
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void write_all(const std::string& path, const std::vector<char>& b) {
    std::ofstream f(path, std::ios::binary);
    if (!f) fail("Error: cannot open output file " + path + ".");
    f.write(b.data(), (std::streamsize)b.size());
    if (!f) fail("Error: cannot write output file " + path + ".");
}

template <typename T>
void put(std::vector<char>& b, T v) {
    const char* p = reinterpret_cast<const char*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
}

}  // namespace

std::string file_extension(const std::string& path) {
    const size_t slash = path.find_last_of("/\\");
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot);
    for (auto& c : e) c = (char)std::toupper((unsigned char)c);
    return e;
}

std::string path_folder(const std::string& path) {
    const size_t slash = path.find_last_of("/\\");
    if (slash == std::string::npos) return ".";
    if (slash == 0) return "/";
    return path.substr(0, slash);
}

// ------------------------------------------------------- parameter file

Config read_config(const std::string& path, const std::function<void(const Config&)>& tables) {
    Config c;
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos)
        fail("Error: filename extension not found in the command-line argument input parameter file");
    const std::string ext = path.substr(dot, 4);
    if (ext == ".xri" || ext == ".XRI") c.xri = true;
    else if (ext == ".txt" || ext == ".TXT") c.xri = false;
    else fail("Error: unknown filename extension in the command-line argument input parameter file");

    ParamReader R(path);
    c.modality = 1;
    if (!c.xri) {  // 1
        auto t = R.next(2, "Error reading units of length from input parameter file.");
        c.modality = (int)inum(t[1]);
    }
    if (c.modality != 0 && c.modality != 1)
        fail("Error: modality and units of length parameter in input parameter file must be 0 for "
             "TEM/Angstroms or 1 for X-ray CT/microns.");
    {  // 2
        auto t = R.next(2, "Error reading verbose output parameter from input parameter file.");
        const long v = inum(t[1]);
        if (v != 0 && v != 1) fail("Error: verbose output parameter must be 0 or 1 in input parameter file.");
        c.verbose = v == 1;
    }
    c.defocus_file = R.next(2, "Error reading file name with rotation angles and defocus distances "
                               "from input parameter file.")[1];  // 3
    c.subseries_file = R.next(2, "Error reading the name of a file with indexes and weights for "
                                 "subseries selection from input parameter file.")[1];  // 4
    if (c.subseries_file != "NONE" && file_extension(c.subseries_file) != ".TXT")
        fail("Error: unrecognised filename extension in the file with indexes and weights for "
             "subseries selection in input parameter file.");
    if (tables) tables(c);
    c.input_template = R.next(2, "Error reading defocus series file name base from input parameter file.")[1];  // 5
    {
        const std::string e = file_extension(c.input_template);
        if (e != ".TIFF" && e != ".TIF" && e != ".RAW" && e != ".GRD")
            fail("Error: input filename extension must be TIF, GRD or RAW.");
    }
    {  // 6
        auto t = R.next(4, "Error reading RAW file parameters from input parameter file.");
        c.raw_header = inum(t[1]);
        c.raw_big_endian = inum(t[2]) != 0;
        c.raw_element = inum(t[3]);
        if (c.raw_element != 4 && c.raw_element != 8)
            fail("Unrecognised ElementLength in input parameter file (only sizeof(float) or "
                 "sizeof(double) are allowed).");
    }
    {  // 7
        auto t = R.next(3, "Error reading input data normalization factors from input parameter file.");
        c.f1 = num(t[1]);
        c.f2 = num(t[2]);
        if (c.f1 == 0) fail("The first normalization factor cannot be zero (leads to division by zero).");
    }
    {  // 8
        auto t = R.next(3, "Error reading the width and height of input images in pixels from input "
                           "parameter file.");
        c.nx = inum(t[1]);
        c.ny = inum(t[2]);
    }
    {  // 9
        auto t = R.next(2, "Error reading pixel size from input parameter file.");
        c.xst = num(t[1]);
        if (c.xst <= 0) fail("Error: pixel size must be positive in input parameter file.");
    }
    c.xlo = -0.5 * c.xst * c.nx;
    c.xhi = c.xlo + c.xst * c.nx;
    c.ylo = -0.5 * c.xst * c.ny;
    c.yhi = c.ylo + c.xst * c.ny;
    c.zlo = c.xlo;
    c.zhi = c.xhi;
    if (!c.xri) {  // 10
        auto t = R.next(4, "Error reading output defocus distances from input parameter file.");
        c.zlo = num(t[1]);
        c.zhi = num(t[2]);
        c.dzextra = num(t[3]);
        if (c.zlo > c.zhi) std::swap(c.zlo, c.zhi);
    }
    c.nz = (long)((c.zhi - c.zlo) / c.xst + 0.5);
    if (c.nz <= 0) fail("Error: number of z steps is not positive.");
    {  // 11
        auto t = R.next(4, "Error reading ccentre of rotation x, y and z shifts from input parameter file.");
        c.dxc = num(t[1]);
        c.dyc = num(t[2]);
        c.dzc = num(t[3]);
    }
    {  // 12
        auto t = R.next(2, "Error reading wavelength from input parameter file.");
        c.wl = num(t[1]);
        if (c.wl <= 0 || c.wl > 1) fail("Error: wavelength value appears to be wrong.");
    }
    if (!c.xri) {  // 13
        auto t = R.next(2, "Error reading objective aperture from input parameter file.");
        c.aobj = num(t[1]);
        if (c.aobj < 0 || c.aobj > 1000) fail("Error: objective aperture value appears to be wrong.");
    }
    if (!c.xri) {  // 14
        auto t = R.next(3, "Error reading spherical aberrations from input parameter file.");
        c.cs3_mm = num(t[1]);
        c.cs5_mm = num(t[2]);
    }
    {  // 15
        auto t = R.next(2, "Error reading CTF correction phase retrieval method parameter from input "
                           "parameter file.");
        c.ctf_mode = (int)inum(t[1]);
        if (c.ctf_mode < 0 || c.ctf_mode > 4)
            fail("Error: unknown value for phase retrieval parameter in input parameter file.");
    }
    {  // 16
        auto t = R.next(2, "Error reading Ewald sphere curvature correction switch from input parameter file.");
        const long v = inum(t[1]);
        if (v != 0 && v != 1)
            fail("Error: unknown value for Ewald sphere curvature correction switch in input parameter file.");
        c.escc = (int)v;
    }
    {  // 17
        auto t = R.next(2, "Error reading absorption coefficient from input parameter file.");
        c.asigma = num(t[1]);
        if (c.xri) {
            if (c.asigma <= 0)
                fail("Error: delta / beta must be positive in the case of .xri parameter file (X-ray imaging mode).");
            c.asigma = -1.0 / c.asigma;
        }
        if (c.modality == 0 && c.asigma < 0) fail("Error: beta/delta must be non-negative in TEM mode.");
        if (c.modality == 1 && c.asigma >= 0)
            fail("Error: beta/delta must be negative in hard X-ray mode (unless .xri parameter file is used).");
        if (c.asigma >= 0 && (c.ctf_mode == 1 || c.ctf_mode == 2))
            fail("Error: beta/delta must be negative in TIE-Hom phase retrieval.");
    }
    {  // 18
        auto t = R.next(2, "Error reading average noise-to-signal ratio in input images from input parameter file.");
        c.nsr = num(t[1]);
        if (c.nsr < 0) fail("Error: NSR in input images must be non-negative.");
    }
    {  // 19
        auto t = R.next(2, "Error reading minimal phase reconstruction error/regulariztion parameter from "
                           "input parameter file.");
        c.epsilon = num(t[1]);
        if (c.epsilon < 0) fail("Error: the regulalization parameter must be non-negative.");
        if (c.epsilon == 0 && c.asigma == 0) fail("Error: when beta/delta == 0, epsilon must be positive.");
    }
    if (!c.xri) {  // 20, 21
        auto t = R.next(2, "Error reading enforcing symmetry mode from input parameter file.");
        c.symmetry_mode = (int)inum(t[1]);
        if (c.symmetry_mode == 1) fail("Error: symmetry enforcing mode has not been implemented in this program yet.");
        if (c.symmetry_mode != 0 && c.symmetry_mode != 2)
            fail("Error: unknown value for symmetry enforcing mode in input parameter file.");
        auto t2 = R.next(2, "Error reading file name with rotation angles enforcing symmetry from input parameter file.");
        c.symmetry_file = t2[1];
        if (c.symmetry_mode != 0 && file_extension(c.symmetry_file) != ".TXT" && !c.symmetry_file.empty())
            fail("Error: filename extension for a file with rotation angles for enforcing symmetry must be "
                 ".txt or the parameter should be empty.");
    }
    {  // 22
        auto t = R.next(2, "Error reading 3D Laplacian filter mode from input parameter file.");
        c.inv_laplace = (int)inum(t[1]);
        if (c.inv_laplace != 0 && c.inv_laplace != 1)
            fail("Error: unknown value for 3D Laplacian filter mode in input parameter file.");
    }
    {  // 23
        auto t = R.next(4, "Error reading low-pass filter width, background subtraction value and lower "
                           "threshold level from input parameter file.");
        c.lp_filter = num(t[1]);
        c.background = num(t[2]);
        c.threshold = num(t[3]);
    }
    if (!c.xri) {  // 24, 25
        auto t = R.next(2, "Error reading peak localization mode from input parameter file.");
        c.peaks = (int)inum(t[1]);
        if (c.peaks != 0 && c.peaks != 1)
            fail("Error: unknown value for the peak localization mode in input parameter file.");
        auto t2 = R.next(3, "Error reading transverse and longitudinal side lengths for peak localization "
                            "from input parameter file.");
        c.atom_xy = num(t2[1]);
        c.atom_z = num(t2[2]);
        if (c.peaks && ((int)(c.atom_xy / c.xst + 0.5) < 2 || (int)(c.atom_z / c.xst + 0.5) < 2))
            fail("Error: transverse and longitudinal side lengths for peak localization must be 2 x z_step or larger.");
    }
    c.output_template = R.next(2, "Error reading output file name base from input parameter file.")[1];  // 26
    const std::string oext = file_extension(c.output_template);
    const bool tiff_out = oext == ".TIFF" || oext == ".TIF", dicom_out = oext == ".DCM";
    if (!tiff_out && !dicom_out && oext != ".GRD")
        fail("Error: output filename extension must be TIF/TIFF, DCM or GRD.");
    {  // 27
        auto t = R.next(4, "Error reading minimum-in, maximum-in and maximum-out for 16-bit output from "
                           "input parameter file.");
        c.b16_minin = (float)num(t[1]);
        c.b16_maxin = (float)num(t[2]);
        c.b16_maxout = inum(t[3]);
        if (tiff_out || dicom_out) {
            if (c.out16()) {
                if (c.b16_minin >= c.b16_maxin || c.b16_maxout <= 0 || c.b16_maxout > 65535)
                    fail("Unsuitable values for minimum-in, maximum-in or maximum-out for 16-bit output from "
                         "input parameter file.");
            } else if (dicom_out) {
                fail("only 16-bit data can be saved in DICOM files.");
            }
        } else if (c.out16()) {
            fail("16-bit output cannot be saved in GRD files.");
        }
    }
    {  // 28
        auto t = R.next(4, "Error reading Euler rotation angles for 3D output from input parameter file.");
        for (int d = 0; d < 3; d++) c.rot_out[d] = num(t[1 + d]);
    }
    {  // 29
        auto t = R.next(2, "Error reading reslicing parameter from input parameter file.");
        c.reslice = (int)inum(t[1]);
        if (c.reslice < 0 || c.reslice > 5)
            fail("Error: unknown value for the output 3D volume reslicing parameter.");
    }
    {  // 30
        auto t = R.next(4, "Error reading axes flipping parameters for 3D output from input parameter file.");
        const char* axis[3] = {"X", "Y", "Z"};
        for (int d = 0; d < 3; d++) {
            const long v = inum(t[1 + d]);
            if (v != 0 && v != 1)
                fail(std::string("Unknown value of ") + axis[d] + " axis flipping parameter in input parameter file.");
            c.flip[d] = v == 1;
        }
    }
    {  // 31
        auto t = R.next(3, "Error reading thick output slices info from input parameter file.");
        c.thick_mode = (int)inum(t[1]);
        c.thick_UL = num(t[2]);
        if (c.thick_mode && (int)(c.thick_UL / c.xst + 0.5) <= 1)
            fail("Slice thickness is smaller than or equal to the voxel thickness.");
        if (c.thick_mode < 0 || c.thick_mode > 2)
            fail("Error: unknown value for thick output slices mode in input parameter file.");
    }
    {  // 32
        auto t = R.next(2, "Error reading save_or_not inverse sampling matrix switch from input parameter file.");
        c.save_sampling = (int)inum(t[1]);
    }
    {  // 33
        auto t = R.next(2, "Error reading import and reprocess existing 3D potential or beta files switch "
                           "from input parameter file.");
        c.reprocess = (int)inum(t[1]);
        if (c.reprocess != 0 && c.reprocess != 1)
            fail("Error: unknown value of import and reprocess existing 3D potential or beta files switch in "
                 "input parameter file.");
        if (c.reprocess == 0) {
            if (c.nx % 2 || c.ny % 2) fail("Error: width and height parameters in input parameter file must be even.");
            if (c.nz % 2) fail("Error: number of z points must be even.");
        }
    }
    {  // 34
        auto t = R.next(2, "Error reading folder name for auxiliary files from input parameter file.");
        c.aux_folder = t[1];
        if (c.aux_folder.back() != '/' && c.aux_folder.back() != '\\') c.aux_folder += "/";
    }
    {  // 35
        auto t = R.next(2, "Error reading number of parallel threads from input parameter file.");
        c.threads = (int)inum(t[1]);
        if (c.threads < 1) fail("Error: the number of parallel threads in input parameter file should be >= 1.");
    }
    return c;
}

// ----------------------------------------------------- orientation files

std::vector<Orientation> read_orientations(const std::string& path) {
    std::ifstream f(path);
    if (!f) fail("Error: cannot open defocus parameter file " + path + ".");
    const std::string ext = file_extension(path);
    const bool relion = ext == ".RELIONNEW";
    if (!relion && ext != ".TXT")
        fail("Error: unrecognised filename extension of the defocus file in the input parameter file.");
    std::vector<Orientation> rows;
    std::string line;
    long lineno = 0;
    while (std::getline(f, line)) {
        lineno++;
        if (starts_comment(line)) continue;
        auto t = split_ws(line);
        Orientation o;
        if (relion) {
            // index, source file, shift x, shift y, defocus x, defocus y, astigmatism angle,
            // rot, tilt, psi; all numbers but the astigmatism angle change sign
            if (t.size() != 10) fail("Error: wrong number of entries in line " + std::to_string(lineno) + " of " + path);
            o.shift_x = -num(t[2]);
            o.shift_y = -num(t[3]);
            const double dfx = -num(t[4]), dfy = -num(t[5]);
            o.astig_dz = 0.5 * (dfx - dfy);
            o.astig_angle = num(t[6]);
            o.angZ = -num(t[7]);
            o.angY = -num(t[8]);
            o.defocus.push_back({-num(t[9]), 0.5 * (dfx + dfy)});
        } else {
            if (t.size() < 4 || t.size() % 2)
                fail("Error: line " + std::to_string(lineno) + " of " + path +
                     " must hold the Z and Y' angles followed by (Z'' angle, defocus) pairs.");
            o.angZ = num(t[0]);
            o.angY = num(t[1]);
            for (size_t k = 2; k + 1 < t.size(); k += 2) o.defocus.push_back({num(t[k]), num(t[k + 1])});
        }
        rows.push_back(o);
    }
    if (rows.empty()) fail("Error: no orientations found in " + path);
    return rows;
}

void read_subseries(const std::string& path, std::vector<int>& idx, std::vector<float>& w) {
    std::ifstream f(path);
    if (!f) fail("Error: cannot open subseries selection file " + path + ".");
    std::string line;
    idx.clear();
    w.clear();
    while (std::getline(f, line)) {
        if (starts_comment(line)) continue;
        auto t = split_ws(line);
        if (t.size() < 2) continue;
        idx.push_back((int)std::lround(num(t[0])));
        w.push_back((float)num(t[1]));
    }
}

// ------------------------------------------------------------ file names

std::vector<std::string> file_names(const std::string& tmpl, long count) {
    const size_t slash = tmpl.find_last_of("/\\");
    size_t dot = tmpl.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) dot = tmpl.size();
    const size_t name0 = slash == std::string::npos ? 0 : slash + 1;
    size_t d0 = dot;
    while (d0 > name0 && std::isdigit((unsigned char)tmpl[d0 - 1])) d0--;
    long start = 0;
    int width;
    if (d0 < dot) {
        start = std::atol(tmpl.substr(d0, dot - d0).c_str());
        width = (int)(dot - d0);
    } else {
        width = (int)std::to_string(count > 1 ? count - 1 : 0).size();
    }
    const std::string head = tmpl.substr(0, d0), tail = tmpl.substr(dot);
    std::vector<std::string> out;
    out.reserve(count);
    for (long i = 0; i < count; i++) {
        std::string n = std::to_string(start + i);
        if ((int)n.size() < width) n = std::string(width - n.size(), '0') + n;
        out.push_back(head + n + tail);
    }
    return out;
}

// ---------------------------------------------------------------- images

Image read_grd(const std::string& path) {
    auto b = read_all(path);
    Image im;
    if (b.size() >= 56 && std::memcmp(b.data(), "DSBB", 4) == 0) {
        int16_t nx, ny;
        std::memcpy(&nx, b.data() + 4, 2);
        std::memcpy(&ny, b.data() + 6, 2);
        double h[6];
        std::memcpy(h, b.data() + 8, 48);
        im.nx = (uint16_t)nx;
        im.ny = (uint16_t)ny;
        im.xlo = h[0];
        im.xhi = h[1];
        im.ylo = h[2];
        im.yhi = h[3];
        const size_t n = (size_t)im.nx * im.ny;
        if (b.size() < 56 + 4 * n) fail("Error: GRD file " + path + " is shorter than its header says.");
        im.v.resize(n);
        std::memcpy(im.v.data(), b.data() + 56, 4 * n);
        return im;
    }
    if (b.size() >= 4 && std::memcmp(b.data(), "DSAA", 4) == 0) {
        std::istringstream is(std::string(b.begin() + 4, b.end()));
        double zlo, zhi;
        is >> im.nx >> im.ny >> im.xlo >> im.xhi >> im.ylo >> im.yhi >> zlo >> zhi;
        if (!is || im.nx <= 0 || im.ny <= 0) fail("Error: bad header in GRD file " + path);
        im.v.resize((size_t)im.nx * im.ny);
        for (auto& x : im.v)
            if (!(is >> x)) fail("Error: GRD file " + path + " is shorter than its header says.");
        return im;
    }
    fail("Error: " + path + " is not a GRD file (no DSBB/DSAA signature).");
}

void write_grd(const std::string& path, long nx, long ny, const float* v, double xlo, double xhi,
               double ylo, double yhi) {
    if (nx > 32767 || ny > 32767) fail("Error: GRD files hold at most 32767 x 32767 values.");
    std::vector<char> b;
    b.reserve(56 + 4 * (size_t)nx * ny);
    b.insert(b.end(), {'D', 'S', 'B', 'B'});
    put<int16_t>(b, (int16_t)nx);
    put<int16_t>(b, (int16_t)ny);
    float mn = 0, mx = 0;
    if (nx * ny > 0) {
        mn = mx = v[0];
        for (size_t i = 1; i < (size_t)nx * ny; i++) {
            mn = std::min(mn, v[i]);
            mx = std::max(mx, v[i]);
        }
    }
    for (double d : {xlo, xhi, ylo, yhi, (double)mn, (double)mx}) put<double>(b, d);
    const char* p = reinterpret_cast<const char*>(v);
    b.insert(b.end(), p, p + 4 * (size_t)nx * ny);
    write_all(path, b);
}

Image read_raw(const std::string& path, long nx, long ny, long header, long element, bool big_endian) {
    // the parameter-file reader checks the element size (UTR.cpp:267-268);
    // the C ABI entry point reaches here directly
    if (element != 4 && element != 8)
        fail("Unrecognised ElementLength in input parameter file (only sizeof(float) or "
             "sizeof(double) are allowed).");
    if (header < 0) fail("Error: RAW header length must not be negative.");
    if (nx <= 0 || ny <= 0) fail("Error: RAW image sizes must be positive.");
    auto b = read_all(path);
    const size_t n = (size_t)nx * ny;
    if (b.size() < (size_t)header + n * element) fail("Error: RAW file " + path + " is too short.");
    Image im;
    im.nx = nx;
    im.ny = ny;
    im.v.resize(n);
    const char* p = b.data() + header;
    for (size_t i = 0; i < n; i++) {
        if (element == 4) {
            uint32_t u;
            std::memcpy(&u, p + 4 * i, 4);
            if (big_endian) u = bswap32(u);
            float f;
            std::memcpy(&f, &u, 4);
            im.v[i] = f;
        } else {
            uint64_t u;
            std::memcpy(&u, p + 8 * i, 8);
            if (big_endian) u = bswap64(u);
            double d;
            std::memcpy(&d, &u, 8);
            im.v[i] = (float)d;
        }
    }
    return im;
}

Image read_tiff(const std::string& path) {
    auto b = read_all(path);
    if (b.size() < 8) fail("Error: " + path + " is not a TIFF file.");
    bool be;
    if (b[0] == 'I' && b[1] == 'I') be = false;
    else if (b[0] == 'M' && b[1] == 'M') be = true;
    else fail("Error: " + path + " is not a TIFF file.");
    auto rd16 = [&](size_t o) -> uint32_t {
        if (o + 2 > b.size()) fail("Error: truncated TIFF file " + path);
        uint16_t v;
        std::memcpy(&v, b.data() + o, 2);
        return be ? bswap16(v) : v;
    };
    auto rd32 = [&](size_t o) -> uint32_t {
        if (o + 4 > b.size()) fail("Error: truncated TIFF file " + path);
        uint32_t v;
        std::memcpy(&v, b.data() + o, 4);
        return be ? bswap32(v) : v;
    };
    if (rd16(2) != 42) fail("Error: " + path + " is not a classic TIFF file.");
    const size_t ifd = rd32(4);
    const uint32_t nent = rd16(ifd);
    uint32_t width = 0, height = 0, bits = 8, comp = 1, spp = 1, fmt = 1, rps = 0;
    std::vector<uint32_t> offs, cnts;
    for (uint32_t e = 0; e < nent; e++) {
        const size_t o = ifd + 2 + 12 * e;
        const uint32_t tag = rd16(o), type = rd16(o + 2), count = rd32(o + 4);
        auto value = [&](uint32_t k) -> uint32_t {
            const size_t sz = type == 3 ? 2 : 4;
            const size_t base = (sz * count <= 4) ? o + 8 : rd32(o + 8);
            return type == 3 ? rd16(base + 2 * k) : rd32(base + 4 * k);
        };
        switch (tag) {
            case 256: width = value(0); break;
            case 257: height = value(0); break;
            case 258: bits = value(0); break;
            case 259: comp = value(0); break;
            case 273: offs.resize(count); for (uint32_t k = 0; k < count; k++) offs[k] = value(k); break;
            case 277: spp = value(0); break;
            case 278: rps = value(0); break;
            case 279: cnts.resize(count); for (uint32_t k = 0; k < count; k++) cnts[k] = value(k); break;
            case 339: fmt = value(0); break;
            default: break;
        }
    }
    (void)rps;
    if (comp != 1) fail("Error: compressed TIFF files are not supported (" + path + ").");
    if (spp != 1) fail("Error: only single-channel TIFF files are supported (" + path + ").");
    if (offs.empty() || offs.size() != cnts.size()) fail("Error: TIFF file " + path + " has no strips.");
    std::vector<char> data;
    for (size_t k = 0; k < offs.size(); k++) {
        if ((size_t)offs[k] + cnts[k] > b.size()) fail("Error: truncated TIFF file " + path);
        data.insert(data.end(), b.begin() + offs[k], b.begin() + offs[k] + cnts[k]);
    }
    const size_t n = (size_t)width * height, bpp = bits / 8;
    if (data.size() < n * bpp) fail("Error: TIFF file " + path + " holds fewer pixels than its size says.");
    Image im;
    im.nx = width;
    im.ny = height;
    im.v.resize(n);
    for (size_t i = 0; i < n; i++) {
        const char* p = data.data() + i * bpp;
        double x;
        if (bits == 8) {
            x = fmt == 2 ? (double)(int8_t)p[0] : (double)(uint8_t)p[0];
        } else if (bits == 16) {
            uint16_t u;
            std::memcpy(&u, p, 2);
            if (be) u = bswap16(u);
            x = fmt == 2 ? (double)(int16_t)u : (double)u;
        } else if (bits == 32) {
            uint32_t u;
            std::memcpy(&u, p, 4);
            if (be) u = bswap32(u);
            if (fmt == 3) {
                float f;
                std::memcpy(&f, &u, 4);
                x = f;
            } else {
                x = fmt == 2 ? (double)(int32_t)u : (double)u;
            }
        } else if (bits == 64 && fmt == 3) {
            uint64_t u;
            std::memcpy(&u, p, 8);
            if (be) u = bswap64(u);
            double d;
            std::memcpy(&d, &u, 8);
            x = d;
        } else {
            fail("Error: unsupported TIFF sample size in " + path);
        }
        im.v[i] = (float)x;
    }
    return im;
}

namespace {
void write_tiff(const std::string& path, long nx, long ny, const void* pix, int bits, int fmt) {
    const uint32_t nbytes = (uint32_t)((size_t)nx * ny * (bits / 8));
    const uint16_t nent = 11;
    const uint32_t ifd = 8, data = ifd + 2 + 12 * nent + 4;
    std::vector<char> b;
    b.insert(b.end(), {'I', 'I'});
    put<uint16_t>(b, 42);
    put<uint32_t>(b, ifd);
    put<uint16_t>(b, nent);
    auto ent = [&](uint16_t tag, uint16_t type, uint32_t val) {
        put<uint16_t>(b, tag);
        put<uint16_t>(b, type);
        put<uint32_t>(b, 1);
        if (type == 3) {
            put<uint16_t>(b, (uint16_t)val);
            put<uint16_t>(b, 0);
        } else {
            put<uint32_t>(b, val);
        }
    };
    ent(256, 4, (uint32_t)nx);     // ImageWidth
    ent(257, 4, (uint32_t)ny);     // ImageLength
    ent(258, 3, (uint32_t)bits);   // BitsPerSample
    ent(259, 3, 1);                // Compression: none
    ent(262, 3, 1);                // Photometric: BlackIsZero
    ent(273, 4, data);             // StripOffsets
    ent(277, 3, 1);                // SamplesPerPixel
    ent(278, 4, (uint32_t)ny);     // RowsPerStrip
    ent(279, 4, nbytes);           // StripByteCounts
    ent(284, 3, 1);                // PlanarConfiguration
    ent(339, 3, (uint32_t)fmt);    // SampleFormat
    put<uint32_t>(b, 0);           // no further IFD
    const char* p = static_cast<const char*>(pix);
    b.insert(b.end(), p, p + nbytes);
    write_all(path, b);
}
}  // namespace

void write_tiff_f32(const std::string& path, long nx, long ny, const float* v) {
    write_tiff(path, nx, ny, v, 32, 3);
}

void write_tiff_u16(const std::string& path, long nx, long ny, const uint16_t* v) {
    write_tiff(path, nx, ny, v, 16, 1);
}

// Minimal DICOM Part-10 file, explicit VR little endian, one 16-bit
// MONOCHROME2 frame (Secondary Capture), with the tags of UTR_DICOM.txt.
void write_dicom_u16(const std::string& path, long nx, long ny, const uint16_t* v,
                     const std::map<std::string, std::string>& tags, long instance,
                     double pixel_spacing_mm) {
    auto tag_or = [&](const char* k, const char* d) {
        auto it = tags.find(k);
        return it == tags.end() ? std::string(d) : it->second;
    };
    auto elem = [](std::vector<char>& b, uint16_t g, uint16_t e, const char* vr, std::string val) {
        const bool ui = std::strcmp(vr, "UI") == 0;
        if (val.size() % 2) val.push_back(ui ? '\0' : ' ');
        put<uint16_t>(b, g);
        put<uint16_t>(b, e);
        b.insert(b.end(), vr, vr + 2);
        const bool longlen = !std::strcmp(vr, "OB") || !std::strcmp(vr, "OW") || !std::strcmp(vr, "UN");
        if (longlen) {
            put<uint16_t>(b, 0);
            put<uint32_t>(b, (uint32_t)val.size());
        } else {
            put<uint16_t>(b, (uint16_t)val.size());
        }
        b.insert(b.end(), val.begin(), val.end());
    };
    auto us = [&](std::vector<char>& b, uint16_t g, uint16_t e, uint16_t x) {
        std::string s(2, '\0');
        std::memcpy(&s[0], &x, 2);
        elem(b, g, e, "US", s);
    };
    // deterministic UIDs under the 2.25 (UUID-derived) root
    std::hash<std::string> H;
    const std::string study = "2.25." + std::to_string(H("study:" + tag_or("StudyInstanceUID", "UTR")) % 100000000000000ull);
    const std::string series = "2.25." + std::to_string(H("series:" + tag_or("SeriesInstanceUID", "UTR")) % 100000000000000ull);
    const std::string inst = series + "." + std::to_string(instance + 1);
    const std::string sop_class = "1.2.840.10008.5.1.4.1.1.7";
    const std::string xfer = "1.2.840.10008.1.2.1";

    std::vector<char> meta;
    elem(meta, 0x0002, 0x0001, "OB", std::string("\x00\x01", 2));
    elem(meta, 0x0002, 0x0002, "UI", sop_class);
    elem(meta, 0x0002, 0x0003, "UI", inst);
    elem(meta, 0x0002, 0x0010, "UI", xfer);
    elem(meta, 0x0002, 0x0012, "UI", "2.25.1");
    std::vector<char> b(128, '\0');
    b.insert(b.end(), {'D', 'I', 'C', 'M'});
    {
        std::string len(4, '\0');
        const uint32_t n = (uint32_t)meta.size();
        std::memcpy(&len[0], &n, 4);
        elem(b, 0x0002, 0x0000, "UL", len);
    }
    b.insert(b.end(), meta.begin(), meta.end());
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g\\%g", pixel_spacing_mm, pixel_spacing_mm);
    elem(b, 0x0008, 0x0016, "UI", sop_class);
    elem(b, 0x0008, 0x0018, "UI", inst);
    elem(b, 0x0008, 0x0050, "SH", tag_or("AccessionNumber", "UTR").substr(0, 16));
    elem(b, 0x0008, 0x0060, "CS", "OT");
    elem(b, 0x0010, 0x0010, "PN", tag_or("PatientName", "UTR"));
    elem(b, 0x0010, 0x0020, "LO", tag_or("PatientID", "UTR"));
    elem(b, 0x0010, 0x0030, "DA", tag_or("PatientBirthDate", "19800101"));
    elem(b, 0x0010, 0x0040, "CS", tag_or("PatientSex", "O") == "Female" ? "F" : (tag_or("PatientSex", "O") == "Male" ? "M" : "O"));
    elem(b, 0x0018, 0x2010, "DS", tag_or("NominalScannedPixelSpacing", "0.1"));
    elem(b, 0x0020, 0x000D, "UI", study);
    elem(b, 0x0020, 0x000E, "UI", series);
    elem(b, 0x0020, 0x0010, "SH", tag_or("StudyID", "UTR").substr(0, 16));
    elem(b, 0x0020, 0x0013, "IS", std::to_string(instance + 1));
    us(b, 0x0028, 0x0002, 1);
    elem(b, 0x0028, 0x0004, "CS", "MONOCHROME2");
    us(b, 0x0028, 0x0010, (uint16_t)ny);
    us(b, 0x0028, 0x0011, (uint16_t)nx);
    elem(b, 0x0028, 0x0030, "DS", buf);
    us(b, 0x0028, 0x0100, 16);
    us(b, 0x0028, 0x0101, 16);
    us(b, 0x0028, 0x0102, 15);
    us(b, 0x0028, 0x0103, 0);
    elem(b, 0x0028, 0x1050, "DS", tag_or("WindowCenter", "2047"));
    elem(b, 0x0028, 0x1051, "DS", tag_or("WindowWidth", "4095"));
    elem(b, 0x7FE0, 0x0010, "OW", std::string(reinterpret_cast<const char*>(v), 2 * (size_t)nx * ny));
    write_all(path, b);
}

std::map<std::string, std::string> dicom_tags(const std::string& path, bool* found) {
    std::map<std::string, std::string> m = {
        {"PatientName", "UTR"},        {"PatientID", "UTR"},
        {"PatientBirthDate", "19800101"}, {"PatientSex", "Female"},
        {"StudyID", "UTR"},            {"AccessionNumber", "UTR_coronal"},
        {"StudyInstanceUID", "UTR"},   {"SeriesInstanceUID", "UTR_coronal"},
        {"NominalScannedPixelSpacing", "0.1"}, {"WindowWidth", "4095"},
        {"WindowCenter", "2047"}};
    std::ifstream f(path);
    if (found) *found = (bool)f;
    if (!f) return m;
    static const char* order[] = {"PatientName", "PatientID", "PatientBirthDate", "PatientSex",
                                  "StudyID", "AccessionNumber", "StudyInstanceUID",
                                  "SeriesInstanceUID", "NominalScannedPixelSpacing",
                                  "WindowWidth", "WindowCenter"};
    std::string line;
    int k = 0;
    while (k < 11 && std::getline(f, line)) {
        if (starts_comment(line)) continue;
        auto t = split_ws(line);
        if (t.size() < 2) fail("Error reading input text file.");
        m[order[k++]] = t[1];
    }
    return m;
}

int pad_or_trim(Image& im, long nx, long ny, float pad_value) {
    const long nx0 = im.nx, ny0 = im.ny;
    if (nx0 == nx && ny0 == ny) return 0;
    std::vector<float> out((size_t)nx * ny);
    if ((nx0 < nx && ny0 <= ny) || (ny0 < ny && nx0 <= nx)) {
        const long top = (ny - ny0) / 2, left = (nx - nx0) / 2;
        std::fill(out.begin(), out.end(), pad_value);
        for (long j = 0; j < ny0; j++)
            std::copy(im.v.begin() + j * nx0, im.v.begin() + (j + 1) * nx0,
                      out.begin() + (j + top) * nx + left);
        im.v.swap(out);
        im.nx = nx;
        im.ny = ny;
        return 1;
    }
    if ((nx0 > nx && ny0 >= ny) || (ny0 > ny && nx0 >= nx)) {
        const long top = (ny0 - ny) / 2, left = (nx0 - nx) / 2;
        for (long j = 0; j < ny; j++)
            std::copy(im.v.begin() + (j + top) * nx0 + left, im.v.begin() + (j + top) * nx0 + left + nx,
                      out.begin() + j * nx);
        im.v.swap(out);
        im.nx = nx;
        im.ny = ny;
        return -1;
    }
    fail("Error: the dimensions of the 2D array in the image file are inconsistent with the array "
         "dimensions in input parameter file (both dimensions should be equal, smaller or larger "
         "simultaneously).");
}

void to_u16(const float* v, size_t n, float minin, float maxin, long maxout, uint16_t* out) {
    const double scale = (double)maxout / ((double)maxin - (double)minin);
    for (size_t i = 0; i < n; i++) {
        const double x = std::min(std::max((double)v[i], (double)minin), (double)maxin);
        out[i] = (uint16_t)std::lround((x - minin) * scale);
    }
}

// ------------------------------------------------------ post-processing

void threshold_low(Volume& vol, float a, float b) {
    for (auto& x : vol.v) {
        x -= a;
        if (x < b) x = b;
    }
}

Volume reslice(const Volume& vol, int mode, bool fx, bool fy, bool fz) {
    // output slices: mode 0 XY, 1 XZ, 2 YZ, 3 YX, 4 ZX, 5 ZY (first letter
    // along the rows of a slice, second down the slice, slices stacked along
    // the remaining axis); flips reverse an input axis first
    const long n[3] = {vol.nx, vol.ny, vol.nz};
    static const int axes[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}, {2, 1, 0}};
    const int* a = axes[mode];
    Volume out;
    out.nx = n[a[0]];
    out.ny = n[a[1]];
    out.nz = n[a[2]];
    out.v.resize(vol.v.size());
    for (long k = 0; k < out.nz; k++)
        for (long j = 0; j < out.ny; j++)
            for (long i = 0; i < out.nx; i++) {
                long src[3];
                src[a[0]] = i;
                src[a[1]] = j;
                src[a[2]] = k;
                if (fx) src[0] = vol.nx - 1 - src[0];
                if (fy) src[1] = vol.ny - 1 - src[1];
                if (fz) src[2] = vol.nz - 1 - src[2];
                out.at(k, j, i) = vol.at(src[2], src[1], src[0]);
            }
    return out;
}

Volume thick_slices(const Volume& vol, int mode, int step) {
    // UTR.cpp:1457-1490: MIP (1) or average (2) over `step` consecutive slices
    if (step > vol.nz) fail("Slice thickness is larger than the sample thickness.");
    Volume out;
    out.nx = vol.nx;
    out.ny = vol.ny;
    out.nz = vol.nz / step;
    out.v.resize((size_t)out.nz * out.ny * out.nx);
    for (long j = 0; j < vol.ny; j++)
        for (long i = 0; i < vol.nx; i++)
            for (long k1 = 0; k1 < out.nz; k1++) {
                const long k = k1 * step;
                float t = vol.at(k, j, i);
                if (mode == 1) {
                    for (long m = 1; m < step; m++) t = std::max(t, vol.at(k + m, j, i));
                } else {
                    for (long m = 1; m < step; m++) t += vol.at(k + m, j, i);
                    t /= (float)step;
                }
                out.at(k1, j, i) = t;
            }
    return out;
}

}  // namespace utrio

// ----------------------------------------------------------------- C ABI

#include "utr_io.h"

namespace {
thread_local std::string g_io_err;
int io_fail(const std::exception& e) {
    g_io_err = e.what();
    return -1;
}
}  // namespace

// ------------------------------------------------------------ peak search

namespace utrio {

std::vector<Peak> find_peaks(Volume& vol, double xlo, double xhi, double ylo, double yhi, double zlo,
                             double zhi, double atom_xy, double atom_z) {
    const long nx = vol.nx, ny = vol.ny, nz = vol.nz;
    const double xst = (xhi - xlo) / nx, yst = (yhi - ylo) / ny, zst = (zhi - zlo) / nz;
    // points outside the sphere of mean-square radius of the box are set to
    // the volume minimum, which excludes them (FindPeaks.cpp:38-63)
    const double xR = (xhi - xlo) / 2.0, yR = (yhi - ylo) / 2.0, zR = (zhi - zlo) / 2.0;
    const double R2 = (xR * xR + yR * yR + zR * zR) / 3.0;
    const float vmin = vol.v.empty() ? 0.f : *std::min_element(vol.v.begin(), vol.v.end());
    for (long k = 0; k < nz; k++) {
        double zz = -zR + zst * k;
        zz *= zz;
        for (long j = 0; j < ny; j++) {
            double yy = -yR + yst * j;
            yy = yy * yy + zz;
            for (long i = 0; i < nx; i++) {
                double xx = -xR + xst * i;
                xx = xx * xx + yy;
                if (xx > R2) vol.at(k, j, i) = vmin;
            }
        }
    }
    // the largest positive value of every atom-sized box, boxes tiled from one
    // box in from the low edges to two boxes from the high edges; everything
    // else in a box is zeroed (FindPeaks.cpp:65-100).  The reference zeroes
    // the previous running maximum, starting from [0][0][0]; that voxel never
    // lies in a box and is overwritten below, so it is left out here.
    const int katom = (int)(atom_z / zst + 0.5), jatom = (int)(atom_xy / yst + 0.5),
              iatom = (int)(atom_xy / xst + 0.5);
    std::vector<Peak> found;
    if (katom <= 0 || jatom <= 0 || iatom <= 0) return found;
    for (long k = katom; k < nz - katom * 2; k += katom)
        for (long j = jatom; j < ny - jatom * 2; j += jatom)
            for (long i = iatom; i < nx - iatom * 2; i += iatom) {
                double vmax = 0.0;
                long km = -1, jm = -1, im = -1;
                for (long kk = k; kk < k + katom; kk++)
                    for (long jj = j; jj < j + jatom; jj++)
                        for (long ii = i; ii < i + iatom; ii++) {
                            float& x = vol.at(kk, jj, ii);
                            if (x > vmax) {
                                vmax = x;
                                if (km >= 0) vol.at(km, jm, im) = 0.f;
                                km = kk;
                                jm = jj;
                                im = ii;
                            } else {
                                x = 0.f;
                            }
                        }
                if (vmax > 0) {
                    Peak p;
                    p.i = (int)im;
                    p.j = (int)jm;
                    p.k = (int)km;
                    p.x = (float)(xlo + xst * im);
                    p.y = (float)(ylo + yst * jm);
                    p.z = (float)(zlo + zst * km);
                    p.height = vol.at(km, jm, im);
                    found.push_back(p);
                }
            }
    // keep peaks in descending height; drop any smaller peak closer than
    // atom_xy to a kept one (FindPeaks.cpp:102-152).  The reference sorts the
    // peaks of its threads with qsort; here they are found in (k, j, i) order
    // and sorted stably by height, ties keeping the later peak first.
    std::vector<int> order(found.size());
    for (size_t n = 0; n < order.size(); n++) order[n] = (int)n;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return found[a].height < found[b].height; });
    std::vector<float> v(order.size());
    for (size_t n = 0; n < order.size(); n++) v[n] = found[order[n]].height;
    const float d2 = (float)(atom_xy * atom_xy);
    std::vector<Peak> kept;
    for (long nn = (long)order.size() - 1; nn >= 0; nn--) {
        if (v[nn] == 0.f) continue;
        const Peak& a = found[order[nn]];
        kept.push_back(a);
        for (long mm = nn - 1; mm >= 0; mm--) {
            if (v[mm] == 0.f) continue;
            const Peak& b = found[order[mm]];
            if ((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z) < d2)
                v[mm] = 0.f;
        }
    }
    // the volume becomes the isolated peaks (FindPeaks.cpp:154-157)
    std::fill(vol.v.begin(), vol.v.end(), 0.f);
    for (const Peak& p : kept) vol.at(p.k, p.j, p.i) = p.height;
    return kept;
}

void write_xyz(const std::string& path, float a, float c, const std::vector<Peak>& peaks) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("Error: cannot open output file " + path + ".");
    std::fprintf(f, "Peak positions in the reconstructed 3D distribution\n");
    std::fprintf(f, "%.7g %.7g %.7g\n", a, a, c);
    for (const Peak& p : peaks)
        std::fprintf(f, "%d %.7g %.7g %.7g %.7g %.7g\n", 6, p.x, p.y, p.z, p.height, 0.0);
    std::fprintf(f, "-1\n");
    if (std::fclose(f) != 0) throw std::runtime_error("Error: cannot write output file " + path + ".");
}

}  // namespace utrio

extern "C" int utr_io_read_image(const char* path, long nx, long ny, long header, long element,
                                 int big_endian, float* out, long cap, long* out_nx, long* out_ny) {
    try {
        const std::string p = path ? path : "";
        const std::string e = utrio::file_extension(p);
        utrio::Image im;
        if (e == ".RAW") im = utrio::read_raw(p, nx, ny, header, element, big_endian != 0);
        else if (e == ".TIF" || e == ".TIFF") im = utrio::read_tiff(p);
        else if (e == ".GRD") im = utrio::read_grd(p);
        else throw std::runtime_error("Error: input filename extension must be TIF, GRD or RAW.");
        if (out_nx) *out_nx = im.nx;
        if (out_ny) *out_ny = im.ny;
        if (out) {
            if ((long)im.v.size() > cap) throw std::runtime_error("Error: output buffer too small for " + p);
            std::copy(im.v.begin(), im.v.end(), out);
        }
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_write_slice(const char* path, long nx, long ny, const float* v, double xlo,
                                  double xhi, double ylo, double yhi, float minin, float maxin,
                                  long maxout) {
    try {
        const std::string p = path ? path : "";
        const std::string e = utrio::file_extension(p);
        const size_t n = (size_t)nx * ny;
        if (e == ".GRD") {
            utrio::write_grd(p, nx, ny, v, xlo, xhi, ylo, yhi);
        } else if (e == ".TIF" || e == ".TIFF" || e == ".DCM") {
            if (maxout > 0) {
                std::vector<uint16_t> u(n);
                utrio::to_u16(v, n, minin, maxin, maxout, u.data());
                if (e == ".DCM")
                    utrio::write_dicom_u16(p, nx, ny, u.data(), utrio::dicom_tags("", nullptr), 0,
                                           (xhi - xlo) / (nx > 0 ? nx : 1) * 1e-3);
                else
                    utrio::write_tiff_u16(p, nx, ny, u.data());
            } else if (e == ".DCM") {
                throw std::runtime_error("only 16-bit data can be saved in DICOM files.");
            } else {
                utrio::write_tiff_f32(p, nx, ny, v);
            }
        } else {
            throw std::runtime_error("Error: output filename extension must be TIF/TIFF, DCM or GRD.");
        }
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_threshold(float* vol, long n, float background, float threshold) {
    try {
        if (!vol || n < 0) throw std::runtime_error("utr_io_threshold: null volume or negative size");
        utrio::Volume v;
        v.nx = n;
        v.ny = v.nz = 1;
        v.v.assign(vol, vol + n);
        utrio::threshold_low(v, background, threshold);
        std::copy(v.v.begin(), v.v.end(), vol);
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_reslice(const float* in, float* out, long nx, long ny, long nz, int mode, int flipx,
                              int flipy, int flipz) {
    try {
        if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0 || mode < 0 || mode > 5)
            throw std::runtime_error("utr_io_reslice: bad arguments");
        utrio::Volume v;
        v.nx = nx;
        v.ny = ny;
        v.nz = nz;
        v.v.assign(in, in + (size_t)nx * ny * nz);
        const utrio::Volume r = utrio::reslice(v, mode, flipx != 0, flipy != 0, flipz != 0);
        std::copy(r.v.begin(), r.v.end(), out);
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_thick_slices(const float* in, float* out, long nx, long ny, long nz, int mode, int step) {
    try {
        if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0 || (mode != 1 && mode != 2) || step < 1)
            throw std::runtime_error("utr_io_thick_slices: bad arguments");
        utrio::Volume v;
        v.nx = nx;
        v.ny = ny;
        v.nz = nz;
        v.v.assign(in, in + (size_t)nx * ny * nz);
        const utrio::Volume r = utrio::thick_slices(v, mode, step);
        std::copy(r.v.begin(), r.v.end(), out);
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_convert16(const float* in, unsigned short* out, long n, float minin, float maxin,
                                long maxout) {
    try {
        if (!in || !out || n < 0) throw std::runtime_error("utr_io_convert16: bad arguments");
        utrio::to_u16(in, (size_t)n, minin, maxin, maxout, out);
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_file_name(const char* tmpl, long count, long index, char* out, long cap) {
    try {
        if (index < 0 || index >= count) throw std::runtime_error("Error: file index out of range.");
        const auto names = utrio::file_names(tmpl ? tmpl : "", count);
        const std::string& s = names[index];
        if ((long)s.size() + 1 > cap) throw std::runtime_error("Error: file name buffer too small.");
        std::copy(s.begin(), s.end(), out);
        out[s.size()] = '\0';
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" int utr_io_find_peaks(float* vol, long nx, long ny, long nz, const double* extent6,
                                 double atom_xy, double atom_z, int* out_ijk, float* out_height,
                                 long cap, long* npeaks) {
    try {
        if (!vol || !extent6 || !npeaks || nx <= 0 || ny <= 0 || nz <= 0)
            throw std::runtime_error("Error: invalid arguments to utr_io_find_peaks.");
        utrio::Volume v;
        v.nx = nx;
        v.ny = ny;
        v.nz = nz;
        v.v.assign(vol, vol + (size_t)nx * ny * nz);
        const auto peaks = utrio::find_peaks(v, extent6[0], extent6[1], extent6[2], extent6[3],
                                             extent6[4], extent6[5], atom_xy, atom_z);
        std::copy(v.v.begin(), v.v.end(), vol);
        *npeaks = (long)peaks.size();
        for (long n = 0; n < (long)peaks.size() && n < cap; n++) {
            if (out_ijk) {
                out_ijk[3 * n] = peaks[n].i;
                out_ijk[3 * n + 1] = peaks[n].j;
                out_ijk[3 * n + 2] = peaks[n].k;
            }
            if (out_height) out_height[n] = peaks[n].height;
        }
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

// ------------------------------------------------------- atom matching

namespace {

// Minimum-cost assignment of the rows of c (n x m, n <= m) to distinct
// columns: shortest augmenting paths with row and column potentials (Kuhn
// and Munkres' Hungarian method in its O(n^2 m) form).  Returns col[i].
std::vector<int> min_cost_assignment(const std::vector<double>& c, int n, int m) {
    const double INF = std::numeric_limits<double>::infinity();
    std::vector<double> u(n + 1, 0.0), v(m + 1, 0.0);
    std::vector<int> p(m + 1, 0), way(m + 1, 0);
    for (int i = 1; i <= n; i++) {
        p[0] = i;
        int j0 = 0;
        std::vector<double> minv(m + 1, INF);
        std::vector<char> used(m + 1, 0);
        do {
            used[j0] = 1;
            const int i0 = p[j0];
            double delta = INF;
            int j1 = 0;
            for (int j = 1; j <= m; j++) {
                if (used[j]) continue;
                const double cur = c[(size_t)(i0 - 1) * m + (j - 1)] - u[i0] - v[j];
                if (cur < minv[j]) {
                    minv[j] = cur;
                    way[j] = j0;
                }
                if (minv[j] < delta) {
                    delta = minv[j];
                    j1 = j;
                }
            }
            for (int j = 0; j <= m; j++) {
                if (used[j]) {
                    u[p[j]] += delta;
                    v[j] -= delta;
                } else {
                    minv[j] -= delta;
                }
            }
            j0 = j1;
        } while (p[j0] != 0);
        do {
            const int j1 = way[j0];
            p[j0] = p[j1];
            j0 = j1;
        } while (j0);
    }
    std::vector<int> col(n, -1);
    for (int j = 1; j <= m; j++)
        if (p[j]) col[p[j] - 1] = j - 1;
    return col;
}

}  // namespace

extern "C" int utr_io_match_atoms(int ntest, const double* test_xyz, const int* test_z, int nref,
                                  const double* ref_xyz, const int* ref_z, double max_dist, int hungarian,
                                  int* match, double* dist, double* stats) {
    try {
        if (ntest < 2 || nref < 2 || !test_xyz || !test_z || !ref_xyz || !ref_z || !match || !dist || !stats)
            throw std::runtime_error("Error: fewer than 2 atoms or null arrays in utr_io_match_atoms.");
        if (!(max_dist > 0)) throw std::runtime_error("Error: maximum match distance must be positive.");
        for (int i = 0; i < nref; i++)
            if (ref_z[i] == 1)
                throw std::runtime_error("Error: hydrogen atom found in the reference data - matching of "
                                         "hydrogen atoms in not supported.");
        // both lists in descending atomic number, ties in input order
        // (pdb-compare.cpp:262-264)
        std::vector<int> ot(ntest), orf(nref);
        for (int i = 0; i < ntest; i++) ot[i] = i;
        for (int i = 0; i < nref; i++) orf[i] = i;
        std::stable_sort(ot.begin(), ot.end(), [&](int a, int b) { return test_z[a] > test_z[b]; });
        std::stable_sort(orf.begin(), orf.end(), [&](int a, int b) { return ref_z[a] > ref_z[b]; });
        // squared distances reference x test (pdb-compare.cpp:363-372)
        std::vector<double> r2((size_t)nref * ntest);
        for (int i = 0; i < nref; i++)
            for (int j = 0; j < ntest; j++) {
                const double* a = ref_xyz + 3 * (size_t)orf[i];
                const double* b = test_xyz + 3 * (size_t)ot[j];
                const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
                r2[(size_t)i * ntest + j] = dx * dx + dy * dy + dz * dz;
            }
        std::vector<int> m(nref, -1);
        if (hungarian) {
            // optimal one-to-one matching of the squared distances; every
            // reference atom is matched when there are enough test atoms
            // (pdb-compare.cpp:373-383: no distance limit in this mode)
            if (nref <= ntest) {
                m = min_cost_assignment(r2, nref, ntest);
            } else {
                std::vector<double> t((size_t)ntest * nref);
                for (int i = 0; i < nref; i++)
                    for (int j = 0; j < ntest; j++) t[(size_t)j * nref + i] = r2[(size_t)i * ntest + j];
                const std::vector<int> mt = min_cost_assignment(t, ntest, nref);
                for (int j = 0; j < ntest; j++)
                    if (mt[j] >= 0) m[mt[j]] = j;
            }
        } else {
            // each reference atom in turn takes the closest unmatched test
            // atom if it is nearer than max_dist (pdb-compare.cpp:384-404)
            std::vector<char> taken(ntest, 0);
            const double d2max = max_dist * max_dist;
            for (int i = 0; i < nref; i++) {
                int jmin = -1;
                double best = std::numeric_limits<double>::max();
                for (int j = 0; j < ntest; j++)
                    if (!taken[j] && r2[(size_t)i * ntest + j] < best) {
                        best = r2[(size_t)i * ntest + j];
                        jmin = j;
                    }
                if (jmin >= 0 && best < d2max) {
                    m[i] = jmin;
                    taken[jmin] = 1;
                }
            }
        }
        // statistics (pdb-compare.cpp:405-433)
        int nmatch = 0;
        double mean = 0.0, dmax = 0.0;
        for (int i = 0; i < nref; i++) {
            match[orf[i]] = m[i] >= 0 ? ot[m[i]] : -1;
            dist[orf[i]] = m[i] >= 0 ? std::sqrt(r2[(size_t)i * ntest + m[i]]) : -1.0;
            if (m[i] < 0) continue;
            const double d = std::sqrt(r2[(size_t)i * ntest + m[i]]);
            nmatch++;
            mean += d;
            dmax = std::max(dmax, d);
        }
        mean = nmatch ? mean / nmatch : 0.0;
        double sd = 0.0;
        for (int i = 0; i < nref; i++)
            if (m[i] >= 0) {
                const double d = std::sqrt(r2[(size_t)i * ntest + m[i]]) - mean;
                sd += d * d;
            }
        sd = nmatch <= 1 ? 0.0 : std::sqrt(sd / (nmatch - 1));
        stats[0] = nmatch;
        stats[1] = nref - nmatch;   // false negatives
        stats[2] = ntest - nmatch;  // false positives
        stats[3] = mean;
        stats[4] = dmax;
        stats[5] = sd;
        return 0;
    } catch (std::exception& ex) {
        return io_fail(ex);
    }
}

extern "C" const char* utr_io_last_error(void) { return g_io_err.c_str(); }
