// utr_plan.hpp -- geometry, derived constants and per-row parameters of one
// UTR reconstruction (host side, plus the POD blocks the kernels read).
//
// The derivation follows MultisliceCppTest/PhaseRetrieval/UTR.cpp:284-388
// (geometry), :211-238 (orientation flags, mean defocus), :1236-1243 (sphere
// mask) and :1273-1277 (final scale).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/utr_hip.h"

namespace utr {

constexpr double kPi = 3.14159265358979323846;
constexpr double kPi180 = kPi / 180.0;
constexpr double kTargetSNR = 1.414213562373;  // UTR.h:12

// Everything a kernel needs about the volume, passed by value.
struct GeomDev {
    int nx, ny, nz, nxd2, nyd2, nzd2;
    int ctf_mode, escc, inv_laplace, rot_centre_shift;
    int use_f1, use_f2;
    float f1, f2;
    double fxst, fyst, fzst, wl2;
    double xlo, xhi, ylo, yhi, zlo, zhi;
    double wl, asigma, epsilon, k2maxo, cs3, cs5;
    double zout_aver;
    double nsr2;  // NSR * sqrt(nx*ny), UTR.cpp:1052
    // sphere mask and final scale
    long ic, jc, kc, nR2;
    int mask_k, mask_j;  // include (k-kc)^2 / (j-jc)^2 terms
    double mask_dtemp, fnorm;
    // 3D filter selection resolved on the host
    int do_tie3d, do_ctf3d, do_invlap;
    double tie_alpha, tie_norm;  // InverseMLaplacian3D arguments for mode 2
    double ctf_eps;              // epsilon actually used by CTFcorrection3D
    double two_d_alpha, two_d_norm;  // 2D TIE-Hom (mode 1)
    double invlap_eps;               // regularisation of the inverse Laplacian (parameter 22)
    // sCTF of the DHT program: Filt3 filter instead of the noise gate, no
    // sphere mask (DHT.cpp:1350-1416)
    int dht_sctf, no_mask;
    double critical_snr;
    // geometry-only factors of the voxel epilogue, set by derive_epilogue()
    // from the fields above (FP64; the kernels only multiply by them)
    double ep_nsr1;                       // mode 1: sqrt(1 + asigma^2) nsr2
    double ep_dksi2, ep_deta2, ep_dzeta2; // mode 1: pi wl zout_aver / aperture^2 (x, y, z)
    double ep_lapx, ep_lapy, ep_lapz;     // InverseMLaplacian3D: 4 pi^2 / aperture^2 (x aperture xhi - ylo)
    double ep_ix2, ep_iy2, ep_iz2;        // CTFcorrection3D: 1 / aperture^2
    double ep_fac2;                       // pi wl zout_aver
    double ep_atans;                      // atan(asigma)
    double ep_c3, ep_c5;                  // pi wl^3 / 2 cs3, pi wl^5 / 3 cs5
    double d_nxd2, d_nzd2;                // nx/2, nz/2 as doubles (grid centre of the scatter)
};

// Fills the ep_* factors of g from its other fields; called whenever those
// change.
void derive_epilogue(GeomDev& g);

// Per orientation row.  m: grid-unit forward map (rows dx1, dy1, dz1 over
// a = i - nx/2, b = j - ny/2, zz physical).  minv: its inverse.
struct alignas(16) RowParams {
    double m[9];
    double ph[3];      // centre-shift phase per grid offset (radians)
    double zout;       // defocus of this row minus the offset (UL)
    double astig_c, astig_s;  // pi wl dz cos 2phi0, pi wl dz sin 2phi0: astigmatism (mode 3)
    float minv[9];
    float wa, wb, wz;  // bounds of |da|,|db|,|dzz| over the unit infinity ball
    float weight;      // subseries weight (1 if none)
    int y_exact;       // dy1 == ny/2 + b exactly (no rotation about Z)
    int pad[3];
};

// Per orientation row: the in-plane correction of its input image before
// the 2D stage (UTR.cpp:960-977): content rotation by -Z'' about the
// rotation centre and the integer shift of a RELION row.
struct PrepRow {
    double cs, sn;  // cos, sin of the content rotation angle
    int dx, dy;     // content shift in pixels (columns, rows)
    int on;         // 0: the image passes unchanged
    int pad;
};

struct Plan {
    utr_params p;
    GeomDev g;
    int nrows = 0;
    std::vector<double> angZ, angY, angZ2, defocus;
    std::vector<PrepRow> prep;
    std::vector<float> weights;
    std::vector<RowParams> rows;
    double EE = 0, xst = 0, zst = 0;
    double xc = 0, yc = 0, zc = 0;
    int rot_y_only = 1, rot_z_only = 1, rot_z2_only = 1;
    int all_y_exact = 0;  // no row rotates about Z (y-plane gridding applies)
};

// Validates and derives; on failure returns false and sets err.
bool build_plan(const utr_params* p, int nrows, const double* angZ_deg, const double* angY_deg,
                const double* angZ2_deg, const double* defocus, const float* weights, Plan& out,
                std::string& err);

// Grid-unit forward map coefficients of one orientation (same expressions
// as the CPU restatement so the FP64 grid indices agree bit for bit).
void angle_coeffs(const GeomDev& g, double angZ_deg, double angY_deg, double* m9);

}  // namespace utr
