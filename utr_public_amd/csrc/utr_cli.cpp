// utr_cli.cpp -- command-line drop-in for UTR.exe on MI355X.
//
//   utr [UTR.txt | params.xri] [--check] [--fp64 | --gather]
//
// Reads the reference's parameter files (UTR.cpp:55-784), orientation and
// subseries files, GRD / RAW / TIFF images (UTR.cpp:918-947), runs the
// reconstruction through libutr_hip.so (UTR.cpp:840-1279), applies the host
// post-processing that is enabled (UTR.cpp:1383-1499) and writes the slice
// stack in GRD, 32/16-bit TIFF or 16-bit DICOM (UTR.cpp:1501-1547).
//
// --check   parse and validate only; print the derived set-up as JSON
// --fp64    use the plain FP64 form instead of the batched FP32 form
// --gather  use the FP32 per-voxel gather form
//
// Differences from the reference: a failure exits with status 1 (the
// reference returns 0); the interactive prompts of UTR.cpp:741-764,1577 are
// only shown when stdin is a terminal, otherwise the default is taken
// (abort for the two questions).
#include <omp.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include "utr_hip.h"
#include "utr_io.hpp"

using namespace utrio;
using Clock = std::chrono::steady_clock;

namespace {

long ms_since(Clock::time_point t0) {
    return (long)std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - t0).count();
}

void check(int rc) {
    if (rc != UTR_OK) throw std::runtime_error(utr_last_error());
}

bool ask_yes(const char* question) {
    if (!isatty(0)) {
        printf("%s (stdin is not a terminal: no)\n", question);
        return false;
    }
    printf("%s", question);
    fflush(stdout);
    const int ch = getchar();
    return ch == 'y' || ch == 'Y';
}

void print_json(const Config& c, const std::vector<Orientation>& rows, const std::vector<float>& w) {
    printf("\n{\"modality\": %d, \"nx\": %ld, \"ny\": %ld, \"nz\": %ld, \"xst\": %.17g, "
           "\"zlo\": %.17g, \"zhi\": %.17g, \"dzextra\": %.17g, \"dxc\": %.17g, \"dyc\": %.17g, "
           "\"dzc\": %.17g, \"wl\": %.17g, \"aobj_mrad\": %.17g, \"cs3_mm\": %.17g, \"cs5_mm\": %.17g, "
           "\"ctf_mode\": %d, \"escc\": %d, \"asigma\": %.17g, \"nsr\": %.17g, \"epsilon\": %.17g, "
           "\"inv_laplace\": %d, \"norm_f1\": %.17g, \"norm_f2\": %.17g, \"threads\": %d, "
           "\"rows\": %zu, \"weights\": %zu, \"input\": \"%s\", \"output\": \"%s\", \"out16\": %d, "
           "\"reslice\": %d, \"thick_mode\": %d, \"save_sampling\": %d, \"reprocess\": %d, "
           "\"angZ0\": %.17g, \"angY0\": %.17g, \"defocus0\": %.17g}\n",
           c.modality, c.nx, c.ny, c.nz, c.xst, c.zlo, c.zhi, c.dzextra, c.dxc, c.dyc, c.dzc, c.wl,
           c.aobj, c.cs3_mm, c.cs5_mm, c.ctf_mode, c.escc, c.asigma, c.nsr, c.epsilon, c.inv_laplace,
           c.f1, c.f2, c.threads, rows.size(), w.size(), c.input_template.c_str(),
           c.output_template.c_str(), (int)c.out16(), c.reslice, c.thick_mode, c.save_sampling,
           c.reprocess, rows[0].angZ, rows[0].angY, rows[0].defocus[0].second);
}

Image read_image(const Config& c, const std::string& path) {
    const std::string e = file_extension(path);
    if (e == ".RAW") return read_raw(path, c.nx, c.ny, c.raw_header, c.raw_element, c.raw_big_endian);
    if (e == ".TIF" || e == ".TIFF") return read_tiff(path);
    return read_grd(path);
}

void write_stack(const Config& c, const Volume& v, const std::vector<std::string>& names,
                 const double ext[3][2], const int* axes, long& write_ms) {
    const auto t0 = Clock::now();
    const std::string e = file_extension(names[0]);
    const size_t n2 = (size_t)v.nx * v.ny;
    bool found = true;
    std::map<std::string, std::string> tags;
    if (e == ".DCM") {
        tags = dicom_tags("UTR_DICOM.txt", &found);
        if (!found) printf("\nWARNING: unable to open UTR_DICOM.txt file; using default values for DICOM tags!");
    }
    const double spacing_mm = c.modality == 0 ? c.xst * 1e-7 : c.xst * 1e-3;
    // 16-bit output (UTR.cpp:1494-1498), converted on the GPU
    std::vector<uint16_t> v16;
    if (e != ".GRD" && c.out16()) {
        v16.resize(v.v.size());
        check(utr_convert16(v.v.data(), v16.data(), (long long)v.v.size(), (float)c.b16_minin, (float)c.b16_maxin,
                            (long long)c.b16_maxout));
    }
    std::string err;
#pragma omp parallel for schedule(dynamic)
    for (long k = 0; k < v.nz; k++) {
        try {
            const float* s = v.v.data() + k * n2;
            if (e == ".GRD") {
                write_grd(names[k], v.nx, v.ny, s, ext[axes[0]][0], ext[axes[0]][1], ext[axes[1]][0],
                          ext[axes[1]][1]);
            } else if (c.out16()) {
                const uint16_t* u = v16.data() + k * n2;
                if (e == ".DCM") write_dicom_u16(names[k], v.nx, v.ny, u, tags, k, spacing_mm);
                else write_tiff_u16(names[k], v.nx, v.ny, u);
            } else {
                write_tiff_f32(names[k], v.nx, v.ny, s);
            }
        } catch (std::exception& ex) {
#pragma omp critical
            err = ex.what();
        }
    }
    if (!err.empty()) throw std::runtime_error(err);
    write_ms += ms_since(t0);
}

int run(int argc, char** argv) {
    std::string pfile = "UTR.txt";
    bool check_only = false;
    int precision = UTR_FP32;
    int argi = 1;
    if (argc > 1 && argv[1][0] != '-') pfile = argv[argi++];
    for (; argi < argc; argi++) {
        const std::string a = argv[argi];
        if (a == "--check") check_only = true;
        else if (a == "--fp64") precision = UTR_FP64;
        else if (a == "--gather") precision = UTR_FP32_GATHER;
        else throw std::runtime_error("Error: unknown command-line option " + a);
    }
    long t_read = 0, t_write = 0, t_peaks = 0;

    printf("\nReading input parameter file %s ...", pfile.c_str());
    // orientation table and subseries selection (UTR.cpp:113-230), read and
    // checked where the reference does: after parameter 4
    std::vector<Orientation> rows;
    std::vector<float> weights;
    std::vector<int> sel;
    long nfull = 0;
    bool single_defocus = true;
    auto tables = [&](const Config& c4) {
        printf("\nReading defocus parameter file %s ...", c4.defocus_file.c_str());
        rows = read_orientations(c4.defocus_file);
        nfull = (long)rows.size();
        if (c4.subseries_file != "NONE") {
            read_subseries(c4.subseries_file, sel, weights);
            if (sel.empty()) throw std::runtime_error("Error: the number of entries in the subseries selection file is equal to zero.");
            if ((long)sel.size() > nfull)
                throw std::runtime_error("Error: the number of entries in the subseries selection file is larger than the number of entries in the defocus file.");
            double total = 0;
            for (size_t i = 0; i < sel.size(); i++) {
                if (sel[i] > nfull - 1)
                    throw std::runtime_error("Error: some frame indexes in the subseries selection file are larger than the number of entries - 1 in the defocus file.");
                if (sel[i] < 0) throw std::runtime_error("Error: some frame indexes in the subseries selection file are negative.");
                total += weights[i];
            }
            if (total == 0)
                throw std::runtime_error("Error: the sum of weights in the second column of the subseries selection file is equal to zero.");
            std::vector<Orientation> r2;
            for (int i : sel) r2.push_back(rows[i]);
            rows.swap(r2);
            printf("\nThere are %zd entries in the subseries selection file.", rows.size());
        }
        size_t ndefmax = 0;
        for (const auto& o : rows) {
            ndefmax = std::max(ndefmax, o.defocus.size());
            if (o.defocus[0].second != rows[0].defocus[0].second) single_defocus = false;
        }
        if (ndefmax > 1)
            throw std::runtime_error("Error: this program is currently implemented only for defocus parameter files with one defocus distance per orientation.");
    };
    Config c = read_config(pfile, tables);
    const long nrows = (long)rows.size();
    printf("\nUnits of length (UL) are set to %s",
           c.modality == 0 ? "Angstroms (TEM imaging mode will be used)" : "microns (X-ray CT imaging mode will be used)");
    if (!std::filesystem::exists(path_folder(c.output_template)))
        throw std::runtime_error("Error: the specified file folder for 3D potential or beta does not seem to exist.");
    if (!std::filesystem::exists(c.aux_folder))
        throw std::runtime_error("Error: the specified auxiliary file folder does not seem to exist.");
    bool astig = false;
    for (const auto& o : rows) astig |= o.astig_dz != 0 || o.astig_angle != 0;
    if (check_only) {
        print_json(c, rows, weights);
        return 0;
    }

    // symmetry rows of parameter 21 (UTR.cpp:501-506), read before the reconstruction
    std::vector<double> symZ, symY, symZ2;
    if (c.symmetry_mode == 2 && !c.symmetry_file.empty()) {
        printf("\nReading symmetry enforcing file %s ...", c.symmetry_file.c_str());
        for (const auto& o : read_orientations(c.symmetry_file)) {
            symZ.push_back(o.angZ);
            symY.push_back(o.angY);
            symZ2.push_back(o.defocus.empty() ? 0.0 : o.defocus[0].first);
        }
    }
    // UTR.cpp:758-766: astigmatism only enters mode 3; modes 2 and 4 warn
    if (!c.reprocess && (c.ctf_mode == 2 || c.ctf_mode == 4) && (!single_defocus || astig)) {
        printf("\n\nWARNING: different defocus distances and/or non-zero astigmatism have been detected in the input file %s.",
               c.defocus_file.c_str());
        printf("\n3D CTF correction have only been implemented for constant defocus distance and no astigmatism, possibly leading to inaccurate results in this case.");
        if (!ask_yes("\nDo you still want to continue (y = yes, any other symbol = no): "))
            throw std::runtime_error("Error: stopped at the user's request.");
    }

    Volume vol;
    utr_timings tm{};
    int nx = 0, ny = 0, nz = 0;
    long nx0 = 0, ny0 = 0;
    bool padded = false;
    utr_ctx* ctx = nullptr;
    struct Guard {
        utr_ctx* c;
        ~Guard() {
            if (c) utr_destroy(c);
        }
    } guard{nullptr};
    std::string out_template = c.output_template;
    if (c.reprocess) {
        // import the slices written by an earlier run (UTR.cpp:729-749, 1282-1338):
        // file names and format from the output template, padded with 0 or
        // trimmed to the volume's x-y size; outputs go to the auxiliary folder
        // with "R" appended (UTR.cpp:1502-1512)
        std::string tin = c.output_template;
        std::string e = file_extension(tin);
        if (e == ".DCM") {
            printf("\nWARNING: input filename extension (in this mode - it is taken from the output filename template) must be TIF, GRD or RAW.");
            const char* q = "\nPress 't' if you want to substitute TIFF extension for input files, 'g' for GRD, 'r' for RAW or any other symbol to exit: ";
            int ch = 0;
            if (isatty(0)) {
                printf("%s", q);
                fflush(stdout);
                ch = getchar();
            } else {
                printf("%s(stdin is not a terminal)\n", q);
            }
            const char* ext = ch == 't' || ch == 'T' ? ".tif" : ch == 'g' || ch == 'G' ? ".grd"
                            : ch == 'r' || ch == 'R' ? ".raw" : nullptr;
            if (!ext) std::exit(1);
            tin.replace(tin.find_last_of('.'), std::string::npos, ext);
            e = file_extension(tin);
        } else if (e != ".TIF" && e != ".TIFF" && e != ".GRD" && e != ".RAW") {
            throw std::runtime_error("Error: input filename extension (in this mode - it is taken from the output filename template) must be TIF, GRD, DCM or RAW.");
        }
        nx = (int)c.nx;
        ny = (int)c.ny;
        // UTR.cpp:308-329
        double zl = c.zlo, zh = c.zhi;
        if (zl > zh) std::swap(zl, zh);
        const long nzl = (long)((zh - zl) / c.xst + 0.5);
        if (nzl <= 0) throw std::runtime_error("Error: number of z steps is not positive.");
        if (nzl % 2) throw std::runtime_error("Error: number of z points must be even.");
        nz = (int)nzl;
        printf("\n\nReading pre-existing 3D distribution of the electrostatic potential or beta from files ...");
        const auto t0 = Clock::now();
        vol.nx = nx;
        vol.ny = ny;
        vol.nz = nz;
        vol.v.assign((size_t)nx * ny * nz, 0.f);
        const std::vector<std::string> vin = file_names(tin, nz);
        std::string err;
        omp_set_num_threads(c.threads);
#pragma omp parallel for schedule(dynamic)
        for (long k = 0; k < nz; k++) {
            try {
                Image im = read_image(c, vin[k]);
                pad_or_trim(im, c.nx, c.ny, 0.0f);
                std::copy(im.v.begin(), im.v.end(), vol.v.begin() + (size_t)k * nx * ny);
            } catch (std::exception& ex) {
#pragma omp critical
                err = ex.what();
            }
        }
        if (!err.empty()) throw std::runtime_error(err);
        t_read += ms_since(t0);
        std::string base = std::filesystem::path(c.output_template).filename().string();
        base.insert(base.find_last_of('.'), "R");
        out_template = c.aux_folder + base;
        printf("\nOutput file name base for the renormalized 3D potential or beta = %s", out_template.c_str());
    } else {
    // input file names (one image per orientation row)
    std::vector<std::string> names = file_names(c.input_template, nfull);
    if (!sel.empty()) {
        std::vector<std::string> n2;
        for (int i : sel) n2.push_back(names[i]);
        names.swap(n2);
    }
    printf("\nInput defocus series file name base = %s", c.input_template.c_str());

    // library context (UTR.cpp:840-867)
    utr_params P{};
    P.nx = (int)c.nx;
    P.ny = (int)c.ny;
    P.modality = c.modality;
    P.xst = c.xst;
    P.zlo = c.zlo;
    P.zhi = c.zhi;
    P.dzextra = c.dzextra;
    P.dxc = c.dxc;
    P.dyc = c.dyc;
    P.dzc = c.dzc;
    P.wl = c.wl;
    P.aobj_mrad = c.aobj;
    P.cs3_mm = c.cs3_mm;
    P.cs5_mm = c.cs5_mm;
    P.ctf_mode = c.ctf_mode;
    P.escc = c.escc;
    P.asigma = c.asigma;
    P.nsr = c.nsr;
    P.epsilon = c.epsilon;
    P.inv_laplace = c.inv_laplace;
    P.norm_f1 = c.f1;
    P.norm_f2 = c.f2;
    std::vector<double> aZ(nrows), aY(nrows), aZ2(nrows), dz(nrows);
    for (long i = 0; i < nrows; i++) {
        aZ[i] = rows[i].angZ;
        aY[i] = rows[i].angY;
        aZ2[i] = rows[i].defocus[0].first;
        dz[i] = rows[i].defocus[0].second;
    }
    check(utr_create(&P, (int)nrows, aZ.data(), aY.data(), aZ2.data(), dz.data(),
                     weights.empty() ? nullptr : weights.data(), precision, &ctx));
    guard.c = ctx;
    {
        // RELION image shifts, undone before the 2D stage (UTR.cpp:961-969)
        std::vector<double> sx(nrows), sy(nrows);
        bool any = false;
        for (long i = 0; i < nrows; i++) {
            sx[i] = rows[i].shift_x;
            sy[i] = rows[i].shift_y;
            any |= sx[i] != 0 || sy[i] != 0;
        }
        if (any) check(utr_set_image_shifts(ctx, sx.data(), sy.data()));
    }
    if (astig) {
        // astigmatism of .relionnew rows (UTR.cpp:1013, mode 3)
        std::vector<double> adz(nrows), aang(nrows);
        for (long i = 0; i < nrows; i++) {
            adz[i] = rows[i].astig_dz;
            aang[i] = rows[i].astig_angle;
        }
        check(utr_set_astigmatism(ctx, adz.data(), aang.data()));
    }
    check(utr_dims(ctx, &nx, &ny, &nz));
    if (c.save_sampling == 1) check(utr_keep_sampling(ctx, 1));

    // read images in batches on the host threads of parameter 35 and stream them in
    printf("\nPerforming phase retrieval and backpropagation into the 3D volume ...");
    omp_set_num_threads(c.threads);
    const long B = std::max<long>(16, std::min<long>(128, 2 * c.threads));
    const size_t n2 = (size_t)c.nx * c.ny;
    std::vector<float> batch(B * n2);
    for (long r0 = 0; r0 < nrows; r0 += B) {
        const long cnt = std::min(B, nrows - r0);
        const auto t0 = Clock::now();
        std::string err;
        std::vector<long> dims(2 * cnt);
        std::vector<int> how(cnt);
#pragma omp parallel for schedule(dynamic)
        for (long r = 0; r < cnt; r++) {
            try {
                Image im = read_image(c, names[r0 + r]);
                dims[2 * r] = im.nx;
                dims[2 * r + 1] = im.ny;
                how[r] = pad_or_trim(im, c.nx, c.ny, 1.0f);
                std::copy(im.v.begin(), im.v.end(), batch.begin() + r * n2);
            } catch (std::exception& ex) {
#pragma omp critical
                err = ex.what();
            }
        }
        if (!err.empty()) throw std::runtime_error(err);
        for (long r = 0; r < cnt; r++) {
            if (nx0 == 0) {
                nx0 = dims[2 * r];
                ny0 = dims[2 * r + 1];
            }
            if (dims[2 * r + 1] != ny0)
                throw std::runtime_error("Error: y-dimension in the input file is different from that in previous input files.");
            if (dims[2 * r] != nx0)
                throw std::runtime_error("Error: x-dimension in the input file is different from that in previous input files.");
            padded |= how[r] > 0;
        }
        t_read += ms_since(t0);
        check(utr_add_images(ctx, (int)r0, (int)cnt, batch.data()));
    }

    vol.nx = nx;
    vol.ny = ny;
    vol.nz = nz;
    vol.v.resize((size_t)nx * ny * nz);
    check(utr_finalize(ctx, vol.v.data()));
    check(utr_get_timings(ctx, &tm));
    }

    double xlo = c.xlo, xhi = c.xhi, ylo = c.ylo, yhi = c.yhi, zlo = c.zlo, zhi = c.zhi;
    if (c.save_sampling == 1 && ctx) {
        Volume S = vol;
        check(utr_get_sampling(ctx, S.v.data()));
        std::string base = std::filesystem::path(c.output_template).filename().string();
        base.replace(base.find_last_of('.'), std::string::npos, "FM.grd");
        auto fm = file_names(c.aux_folder + base, nz);
        printf("\n\nSaving the sampling matrix in files %s, etc. ...", fm[0].c_str());
        const double ext[3][2] = {{xlo, xhi}, {ylo, yhi}, {zlo, zhi}};
        const int ax[3] = {0, 1, 2};
        Config cg = c;
        cg.b16_minin = cg.b16_maxin = 0;
        cg.b16_maxout = 0;
        write_stack(cg, S, fm, ext, ax, t_write);
    }

    // post-processing (UTR.cpp:1341-1490)
    if (symZ.size() > 1) {
        // known symmetry (UTR.cpp:1341-1355), on the GPU
        printf("\nEnforcing known symmetries of the reconstructed 3D potential or beta ...");
        check(utr_symmetrize3d(vol.v.data(), nx, ny, nz, (int)symZ.size(), symZ.data(), symY.data(),
                               symZ2.data()));
    }
    const double zst_v = (c.zhi - c.zlo) / (double)nz;
    if (c.reprocess && c.inv_laplace) {
        // UTR.cpp:1363-1367: the regularised inverse (-Laplacian) in real space, on the GPU
        printf("\nInverse Laplace filtering the reconstructed 3D object ...");
        check(utr_invlaplace3d(vol.v.data(), nx, ny, nz, c.xst, c.xst, std::fabs(zst_v), c.epsilon, c.epsilon));
    }
    if (c.lp_filter > 0 && c.lp_filter < (c.xhi - c.xlo)) {
        // high-pass: the volume minus its Gaussian low-pass (UTR.cpp:1368-1374), on the GPU
        printf("\nHigh-pass filtering the reconstructed 3D object ...");
        const double zst = (c.zhi - c.zlo) / (double)nz;
        check(utr_highpass3d(vol.v.data(), nx, ny, nz, c.xst, c.xst, zst, c.lp_filter));
    }
    float vmin = vol.v.empty() ? 0.f : *std::min_element(vol.v.begin(), vol.v.end());
    if (c.background != 0 || vmin < c.threshold) {
        printf("\nThresholding the reconstructed 3D object ...");
        check(utr_threshold(vol.v.data(), (long long)vol.v.size(), (float)c.background, (float)c.threshold));
    }
    if (c.rot_out[0] != 0 || c.rot_out[1] != 0 || c.rot_out[2] != 0) {
        // output rotation (UTR.cpp:656-661, 1387-1394): the angles as written, negated
        const double aZ = -c.rot_out[0], aY = -c.rot_out[1], aZ2 = -c.rot_out[2];
        printf("\nRotating around Z by %g (deg), around Y' by %g (deg) and around Z'' by %g (deg) ...", aZ, aY, aZ2);
        check(utr_rotate3d(vol.v.data(), nx, ny, nz, aZ, aY, aZ2));
    }
    if (padded) {
        // trim back to the size of the input images (UTR.cpp:1406-1424)
        const long cx0 = (nx - nx0) / 2, cy0 = (ny - ny0) / 2, cz0 = (nx - nx0) / 2;
        const long nz1 = nz - (nx - nx0);
        if (nz1 <= 0) throw std::runtime_error("Error: the volume is thinner than the padding of the input images.");
        Volume t;
        t.nx = nx0;
        t.ny = ny0;
        t.nz = nz1;
        t.v.resize((size_t)t.nx * t.ny * t.nz);
        for (long k = 0; k < t.nz; k++)
            for (long j = 0; j < t.ny; j++)
                for (long i = 0; i < t.nx; i++) t.at(k, j, i) = vol.at(k + cz0, j + cy0, i + cx0);
        vol.v.swap(t.v);
        vol.nx = t.nx;
        vol.ny = t.ny;
        vol.nz = t.nz;
        xlo = -0.5 * c.xst * nx0;
        xhi = xlo + c.xst * nx0;
        ylo = -0.5 * c.xst * ny0;
        yhi = ylo + c.xst * ny0;
        if (c.modality) {
            zlo = xlo;
            zhi = xhi;
        } else {
            zlo += (double)((nx - nx0) / 2) * c.xst;
            zlo -= (double)(nx - nx0 - (nx - nx0) / 2) * c.xst;
        }
    }
    if (c.modality == 0) {  // positive octant (UTR.cpp:1418-1424)
        xhi -= xlo;
        xlo = 0;
        yhi -= ylo;
        ylo = 0;
        zhi -= zlo;
        zlo = 0;
    }
    if (c.peaks) {
        // peak localization (UTR.cpp:1427-1446, FindPeaks.cpp:14-171): the volume
        // becomes the isolated peaks, listed in an XYZ file beside the output
        // and written as slices into the auxiliary folder with "A" appended
        printf("\nSearching for peak positions in the 3D distribution of the electrostatic potential or beta ...");
        const auto tp = Clock::now();
        const auto peaks = find_peaks(vol, xlo, xhi, ylo, yhi, zlo, zhi, c.atom_xy, c.atom_z);
        if (peaks.empty()) {
            printf("\n\n!!WARNING: no peaks have been found!");
        } else {
            printf("\n%d peak positions have been found in total.", (int)peaks.size());
            constexpr size_t kMaxAtoms = 500000;  // natommax, UTR.h:10
            std::string xyz = c.output_template;
            xyz.replace(xyz.find_last_of('.') + 1, std::string::npos, "xyz");
            if (peaks.size() > kMaxAtoms) {
                printf("\nNot saving the reconstructed peak-localized 3D object into an XYZ file, since too many peaks (%d) have been found.",
                       (int)peaks.size());
            } else {
                printf("\nSaving the reconstructed peak-localized 3D object into an XYZ file %s ...", xyz.c_str());
                write_xyz(xyz, (float)(xhi - xlo), (float)(zhi - zlo), peaks);
            }
        }
        std::string base = std::filesystem::path(c.output_template).filename().string();
        base.insert(base.find_last_of('.'), "A");
        const std::string e = file_extension(base);
        if (e != ".TIF" && e != ".TIFF") base.replace(base.find_last_of('.'), std::string::npos, ".grd");
        auto names = file_names(c.aux_folder + base, vol.nz);
        printf("\nSaving the reconstructed peak-localized 3D object into output files %s, etc. ...", names[0].c_str());
        Config cg = c;
        cg.b16_minin = cg.b16_maxin = 0;
        cg.b16_maxout = 0;
        const double pext[3][2] = {{xlo, xhi}, {ylo, yhi}, {zlo, zhi}};
        const int pax[3] = {0, 1, 2};
        write_stack(cg, vol, names, pext, pax, t_write);
        t_peaks = std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - tp).count();
    }
    static const int axes_of[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}, {2, 1, 0}};
    if (c.reslice || c.flip[0] || c.flip[1] || c.flip[2]) {
        printf("\nReslicing the output 3D volume ...");
        const long n[3] = {vol.nx, vol.ny, vol.nz};
        const int* a = axes_of[c.reslice];
        Volume r;
        r.nx = n[a[0]];
        r.ny = n[a[1]];
        r.nz = n[a[2]];
        r.v.resize(vol.v.size());
        check(utr_reslice(vol.v.data(), r.v.data(), (int)vol.nx, (int)vol.ny, (int)vol.nz, c.reslice, c.flip[0],
                          c.flip[1], c.flip[2]));
        vol = std::move(r);
    }
    if (c.thick_mode > 0) {
        printf("\nCreating thick output slices ...");
        const int step = (int)(c.thick_UL / c.xst + 0.5);
        if (step > vol.nz) throw std::runtime_error("Slice thickness is larger than the sample thickness.");
        Volume t;
        t.nx = vol.nx;
        t.ny = vol.ny;
        t.nz = step > 0 ? vol.nz / step : 0;
        t.v.resize((size_t)t.nx * t.ny * t.nz);
        check(utr_thick_slices(vol.v.data(), t.v.data(), (int)vol.nx, (int)vol.ny, (int)vol.nz, c.thick_mode, step));
        vol = std::move(t);
    }
    const double ext[3][2] = {{xlo, xhi}, {ylo, yhi}, {zlo, zhi}};
    auto outnames = file_names(out_template, vol.nz);
    printf("\n\nSaving the reconstructed 3D object into output files %s, etc. ...", outnames[0].c_str());
    write_stack(c, vol, outnames, ext, axes_of[c.reslice], t_write);

    if (c.verbose) {
        printf("\nTotal file read time = %ld (ms)", t_read);
        printf("\nTotal file write time = %ld  (ms)", t_write);
        printf("\nTotal 2D FFT and CTF correction time = %ld (ms)", (long)std::lround(tm.fft2d_ctf_ms));
        printf("\nTotal backpropagation time = %ld (ms)", (long)std::lround(tm.backprop_ms));
        printf("\nTotal 3D FFT time = %ld (ms)", (long)std::lround(tm.fft3d_ms));
        printf("\nTotal peak localization time = %ld (ms)", t_peaks);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const auto t0 = Clock::now();
    printf("\nStarting Unified Tomographic Reconstruction program ...");
    int rc = 0;
    try {
        rc = run(argc, argv);
    } catch (std::exception& e) {
        printf("\n\n!!!Exception: %s\n", e.what());
        rc = 1;
    }
    printf("\n\nMain program finished. Execution time = %ld s.\n", ms_since(t0) / 1000);
    if (isatty(0)) {
        printf("Press any key to exit...");
        getchar();
    }
    return rc;
}
