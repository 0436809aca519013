// utr_plan.cpp -- validation and derived constants of a UTR reconstruction.
//
// Mirrors the parameter checks and derivations of
// MultisliceCppTest/PhaseRetrieval/UTR.cpp:77-782 that bear on the numeric
// path; error texts are the reference's where it has one.
#include "utr_plan.hpp"

#include <cmath>
#include <cstring>

namespace utr {

void angle_coeffs(const GeomDev& g, double angZ_deg, double angY_deg, double* m) {
    // UTR.cpp:883-888 (angles to radians) and UTR.h:243-256 (rotation about
    // Y' then Z), written in grid units relative to the grid centre:
    // dx1 = nx/2 + a*m0 + b*m1 + zz*m2, etc.
    const double aZ = angZ_deg * kPi180, aY = angY_deg * kPi180;
    const double cY = std::cos(aY), sY = std::sin(aY), cZ = std::cos(aZ), sZ = std::sin(aZ);
    m[0] = cY * cZ;
    m[1] = sZ * (g.fyst / g.fxst);
    m[2] = -(sY * cZ) / g.fxst;
    m[3] = -(cY * sZ) * (g.fxst / g.fyst);
    m[4] = cZ;
    m[5] = (sY * sZ) / g.fyst;
    m[6] = sY * (g.fxst / g.fzst);
    m[7] = 0.0;
    m[8] = cY / g.fzst;
}

static bool invert3(const double* m, double* inv) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (!(std::fabs(det) > 0)) return false;
    const double id = 1.0 / det;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return true;
}

bool build_plan(const utr_params* p, int nrows, const double* angZ, const double* angY,
                const double* angZ2, const double* defocus, const float* weights, Plan& P,
                std::string& err) {
    P = Plan();
    if (!p) { err = "null parameters"; return false; }
    if (nrows < 1) { err = "Error: the orientation file contains no rows."; return false; }
    if (!angZ || !angY || !defocus) { err = "null orientation arrays"; return false; }
    if (p->modality != 0 && p->modality != 1) {
        err = "Error: modality and units of length parameter in input parameter file must be 0 for TEM/Angstroms or 1 for X-ray CT/microns.";
        return false;
    }
    if (p->norm_f1 == 0) {
        err = "The first normalization factor cannot be zero (leads to division by zero).";
        return false;
    }
    if (p->nx <= 0 || p->ny <= 0 || (p->nx % 2) || (p->ny % 2)) {
        err = "Error: width and height parameters in input parameter file must be even.";
        return false;
    }
    if (!(p->xst > 0)) { err = "Error: pixel size must be positive in input parameter file."; return false; }
    if (!(p->wl > 0) || p->wl > 1) { err = "Error: wavelength value appears to be wrong."; return false; }
    if (p->aobj_mrad < 0 || p->aobj_mrad > 1000) {
        err = "Error: objective aperture value appears to be wrong.";
        return false;
    }
    if (p->ctf_mode < 0 || p->ctf_mode > 4) {
        err = "Error: unknown value for phase retrieval parameter in input parameter file.";
        return false;
    }
    if (p->escc != 0 && p->escc != 1) {
        err = "Error: unknown value for Ewald sphere curvature correction switch in input parameter file.";
        return false;
    }
    if (p->modality == 0 && p->asigma < 0) { err = "Error: beta/delta must be non-negative in TEM mode."; return false; }
    if (p->modality == 1 && p->asigma >= 0) {
        err = "Error: beta/delta must be negative in hard X-ray mode (unless .xri parameter file is used).";
        return false;
    }
    if (p->asigma >= 0 && (p->ctf_mode == 1 || p->ctf_mode == 2)) {
        err = "Error: beta/delta must be negative in TIE-Hom phase retrieval.";
        return false;
    }
    if (p->nsr < 0) { err = "Error: NSR in input images must be non-negative."; return false; }
    if (p->epsilon < 0) { err = "Error: the regulalization parameter must be non-negative."; return false; }
    if (p->epsilon == 0 && p->asigma == 0) {
        err = "Error: when beta/delta == 0, epsilon must be positive.";
        return false;
    }
    if (p->inv_laplace != 0 && p->inv_laplace != 1) {
        err = "Error: unknown value for 3D Laplacian filter mode in input parameter file.";
        return false;
    }
    P.p = *p;
    P.nrows = nrows;
    P.angZ.assign(angZ, angZ + nrows);
    P.angY.assign(angY, angY + nrows);
    P.defocus.assign(defocus, defocus + nrows);
    P.angZ2.assign(nrows, 0.0);
    if (angZ2) P.angZ2.assign(angZ2, angZ2 + nrows);
    // UTR.cpp:971-976: Rotate(int0, -Z'') turns the image content by -Z''
    P.prep.assign(nrows, PrepRow{});
    for (int i = 0; i < nrows; i++) {
        PrepRow& q = P.prep[i];
        const double a = -P.angZ2[i] * (3.14159265358979323846 / 180.0);
        q.cs = std::cos(a);
        q.sn = std::sin(a);
        q.on = P.angZ2[i] != 0 ? 1 : 0;
    }
    P.weights.assign(nrows, 1.0f);
    if (weights)
        for (int i = 0; i < nrows; i++) P.weights[i] = weights[i];

    GeomDev& g = P.g;
    std::memset(&g, 0, sizeof g);
    g.nx = p->nx; g.ny = p->ny; g.nxd2 = p->nx / 2; g.nyd2 = p->ny / 2;
    g.ctf_mode = p->ctf_mode; g.escc = p->escc; g.inv_laplace = p->inv_laplace;
    // UTR.cpp:294-306
    const double xst = p->xst, yst = p->xst;
    P.xst = xst;
    g.xlo = -0.5 * xst * g.nx; g.xhi = g.xlo + xst * g.nx;
    g.ylo = -0.5 * yst * g.ny; g.yhi = g.ylo + yst * g.ny;
    g.fxst = 1.0 / (g.xhi - g.xlo); g.fyst = 1.0 / (g.yhi - g.ylo);
    // UTR.cpp:308-329
    double zlo = p->zlo, zhi = p->zhi;
    if (zlo > zhi) std::swap(zlo, zhi);
    g.zlo = zlo; g.zhi = zhi;
    const double zst = xst;
    P.zst = zst;
    const long nz = (long)((zhi - zlo) / zst + 0.5);
    if (nz <= 0) { err = "Error: number of z steps is not positive."; return false; }
    if (nz % 2) { err = "Error: number of z points must be even."; return false; }
    g.nz = (int)nz; g.nzd2 = g.nz / 2;
    g.fzst = 1.0 / (zhi - zlo);
    g.wl = p->wl; g.wl2 = 0.5 * p->wl;
    g.asigma = p->asigma; g.epsilon = p->epsilon;
    // UTR.cpp:232-238,330
    double za = defocus[0];
    for (int i = 1; i < nrows; i++) za += defocus[i];
    za /= nrows;
    g.zout_aver = za - p->dzextra;
    // UTR.cpp:335-345
    g.rot_centre_shift = (p->dxc != 0 || p->dyc != 0 || p->dzc != 0);
    P.xc = (g.xhi + g.xlo) / 2.0 + p->dxc;
    P.yc = (g.yhi + g.ylo) / 2.0 + p->dyc;
    P.zc = (g.zhi + g.zlo) / 2.0 + p->dzc;
    g.ic = (long)((P.xc - g.xlo) / xst + 0.5);
    g.jc = (long)((P.yc - g.ylo) / yst + 0.5);
    g.kc = (long)((P.zc - g.zlo) / zst + 0.5);
    // UTR.cpp:354-370
    if (p->modality == 0) {
        constexpr double hp = 6.62607004e-34, cc = 299792458, ee = 1.602176634e-19,
                         m0 = 9.1093837015e-31;
        constexpr long double mc2 = m0 * cc * cc;
        long double chl2 = (long double)(cc * cc * hp * hp) / (long double)(p->wl * p->wl * 1.e-20);
        long double abra = std::sqrt(mc2 * mc2 + chl2);
        P.EE = (double)((-mc2 + abra) / (long double)(ee));
    } else {
        P.EE = 0.001239841662513396 / p->wl;
    }
    // UTR.cpp:388 -- the 0.001f literal is single precision in the reference
    g.k2maxo = std::pow(p->aobj_mrad * (double)0.001f / p->wl, 2.0);
    g.cs3 = p->cs3_mm * 1.e+7; g.cs5 = p->cs5_mm * 1.e+7;
    g.use_f1 = p->norm_f1 != 1.0; g.use_f2 = p->norm_f2 != 0.0;
    g.f1 = (float)p->norm_f1; g.f2 = (float)p->norm_f2;
    g.nsr2 = p->nsr * std::sqrt((double)((size_t)g.nx * g.ny));
    // UTR.cpp:211-222 (the loop starts at row 1)
    P.rot_y_only = P.rot_z_only = P.rot_z2_only = 1;
    for (int i = 1; i < nrows; i++) {
        if (angZ[i] != 0) { P.rot_y_only = 0; P.rot_z2_only = 0; }
        if (angY[i] != 0) { P.rot_z_only = 0; P.rot_z2_only = 0; }
        if (angZ2 && angZ2[i] != 0) { P.rot_y_only = 0; P.rot_z_only = 0; }
    }
    // UTR.cpp:1236-1248
    const long nxR2 = (g.ic <= g.nxd2) ? g.ic * g.ic : (g.nx - g.ic) * (g.nx - g.ic);
    const long nyR2 = (g.jc <= g.nyd2) ? g.jc * g.jc : (g.ny - g.jc) * (g.ny - g.jc);
    const long nzR2 = (g.kc <= g.nzd2) ? g.kc * g.kc : (g.nz - g.kc) * (g.nz - g.kc);
    if (P.rot_y_only) g.nR2 = std::min(nxR2, nzR2);
    else if (P.rot_z_only || P.rot_z2_only) g.nR2 = std::min(nxR2, nyR2);
    else g.nR2 = std::min(std::min(nxR2, nyR2), nzR2);
    g.mask_dtemp = 2.0 / (double)(g.kc * g.kc + g.jc * g.jc + g.ic * g.ic);
    g.mask_k = !P.rot_z_only && !P.rot_z2_only;
    g.mask_j = !P.rot_y_only;
    // UTR.cpp:1273-1277
    const double s2 = std::sqrt(1.0 + p->asigma * p->asigma);
    if (p->modality == 0) g.fnorm = P.EE * p->wl / (2.0 * kPi * s2 * zst);
    else g.fnorm = p->asigma * p->wl / (4.0 * kPi * s2 * zst);
    // 3D filters, UTR.cpp:1178-1201
    if (g.zout_aver != 0 && g.ctf_mode == 2) {
        g.do_tie3d = 1;
        g.tie_alpha = -4.0 * kPi * p->asigma / (g.zout_aver * p->wl);
        g.tie_norm = -g.tie_alpha * s2 / p->asigma;
        if (g.tie_alpha <= 0) { err = "InverseMLaplacian3D() cannot be called with alpha <= 0"; return false; }
    }
    if (g.zout_aver != 0 && g.ctf_mode == 4) {
        g.do_ctf3d = 1;
        // UTR.h:45-81 (zst from GetZStep(ny), as in the reference)
        const double xs = (g.xhi - g.xlo) / g.nx, ys = (g.yhi - g.ylo) / g.ny,
                     zs = (g.zhi - g.zlo) / g.ny;
        const double srad = std::sqrt(0.25 / (xs * xs) + 0.25 / (ys * ys) + 0.25 / (zs * zs));
        if (srad * p->wl > 1.0) {
            err = "runtime_error in CTFcorrection3D() (evanescent waves present)";
            return false;
        }
        const double dk2 = 1.0 / ((g.xhi - g.xlo) * (g.xhi - g.xlo));
        const double de2 = 1.0 / ((g.yhi - g.ylo) * (g.yhi - g.ylo));
        const double dz2 = 1.0 / ((g.zhi - g.zlo) * (g.zhi - g.zlo));
        const double fac2 = kPi * p->wl * g.zout_aver;
        g.ctf_eps = p->epsilon;
        if (g.ctf_eps <= 0) g.ctf_eps = 0.1 * std::pow(std::fabs(fac2) * std::min(dk2, std::min(de2, dz2)), 2);
    }
    g.invlap_eps = p->epsilon;
    if (g.inv_laplace && g.ctf_mode != 1 && g.ctf_mode != 2) {
        g.do_invlap = 1;
        if (p->epsilon <= 0) { err = "InverseMLaplacian3D() cannot be called with alpha <= 0"; return false; }
    }
    if (g.ctf_mode == 1) {
        g.two_d_alpha = -4.0 * kPi * p->asigma / (g.zout_aver * p->wl);
        g.two_d_norm = -g.two_d_alpha * s2 / p->asigma;
    }

    // per-row parameters
    P.rows.resize(nrows);
    for (int r = 0; r < nrows; r++) {
        RowParams& R = P.rows[r];
        std::memset(&R, 0, sizeof R);
        angle_coeffs(g, angZ[r], angY[r], R.m);
        double inv[9];
        if (!invert3(R.m, inv)) { err = "singular orientation"; return false; }
        for (int t = 0; t < 9; t++) R.minv[t] = (float)inv[t];
        R.wa = (float)(std::fabs(inv[0]) + std::fabs(inv[1]) + std::fabs(inv[2]));
        R.wb = (float)(std::fabs(inv[3]) + std::fabs(inv[4]) + std::fabs(inv[5]));
        R.wz = (float)(std::fabs(inv[6]) + std::fabs(inv[7]) + std::fabs(inv[8]));
        // with no rotation about Z the y grid coordinate is ny/2 + b exactly
        R.y_exact = (R.m[3] == 0.0 && R.m[4] == 1.0 && R.m[5] == 0.0);
        if (g.rot_centre_shift) {
            // UTR.cpp:893-899; the phase of UTR.h:264 per grid offset
            const double aZ = angZ[r] * kPi180, aY = angY[r] * kPi180;
            const double cY = std::cos(aY), sY = std::sin(aY), cZ = std::cos(aZ), sZ = std::sin(aZ);
            const double tPI = 2.0 * kPi;
            const double Ax = tPI * (p->dxc * (1.0 - cY * cZ) - p->dyc * sZ + p->dzc * sY * cZ);
            const double Ay = tPI * (p->dxc * cY * sZ + p->dyc * (1.0 - cZ) - p->dzc * sY * sZ);
            const double Az = tPI * (-p->dxc * sY + p->dzc * (1.0 - cY));
            R.ph[0] = Ax * g.fxst; R.ph[1] = Ay * g.fyst; R.ph[2] = Az * g.fzst;
        }
        R.zout = defocus[r] - p->dzextra;
        R.weight = P.weights[r];
    }
    P.all_y_exact = nrows > 0 && nrows <= 65535;
    for (int r = 0; r < nrows; r++)
        if (!P.rows[r].y_exact) P.all_y_exact = 0;
    derive_epilogue(g);
    return true;
}

void derive_epilogue(GeomDev& g) {
    const double fac2 = kPi * g.wl * g.zout_aver;
    const double xa = g.xhi - g.xlo, ya = g.yhi - g.ylo, za = g.zhi - g.zlo;
    // mode-1 NSR (UTR.cpp:1085-1110)
    g.ep_nsr1 = std::sqrt(1.0 + g.asigma * g.asigma) * g.nsr2;
    g.ep_dksi2 = fac2 / (xa * xa);
    g.ep_deta2 = fac2 / (ya * ya);
    g.ep_dzeta2 = fac2 / (za * za);
    // InverseMLaplacian3D(V3, zhi-zlo, yhi-ylo, xhi-ylo, ...) (UTR.cpp:1184-1196)
    const double fact = 4.0 * kPi * kPi, xl = g.xhi - g.ylo;
    g.ep_lapx = fact / (xl * xl);
    g.ep_lapy = fact / (ya * ya);
    g.ep_lapz = fact / (za * za);
    // CTFcorrection3D (UTR.h:86-131)
    g.ep_ix2 = 1.0 / (xa * xa);
    g.ep_iy2 = 1.0 / (ya * ya);
    g.ep_iz2 = 1.0 / (za * za);
    g.ep_fac2 = fac2;
    g.ep_atans = std::atan(g.asigma);
    g.ep_c3 = kPi * g.wl * g.wl * g.wl / 2.0 * g.cs3;
    g.ep_c5 = kPi * g.wl * g.wl * g.wl * g.wl * g.wl / 3.0 * g.cs5;
    g.d_nxd2 = (double)g.nxd2;
    g.d_nzd2 = (double)g.nzd2;
}

}  // namespace utr
