// evaluate.hip -- evaluation of a reconstruction against the forward model
// on gfx950, behind the utr_ground_truth_device and utr_reproject_device C
// ABI of include/utr_hip.h.  Included by utr_api.hip after msct.hip (one
// translation unit).
//
// Ground truth.  The reference's ground-truth potential is autosliccmd's
// noutput 3 (autosliccmd.cpp:1146-1151): vatom (slicelib.cpp:2699-2737)
// summed over all atoms at each pixel, a point sample of a function with a
// Coulomb singularity at every nucleus (r clamped at 1e-10 A).  Compared
// voxel by voxel with a reconstruction on a grid, that sum is dominated by
// whichever voxels fall near a nucleus.  The volume here is the same
// potential restricted to the spatial frequencies below qmax, evaluated
// exactly on the periodic cube of side L = n xst by Fourier synthesis:
//   V(r) = 1/L^3 sum_{|k/L| < qmax} Vt(|k|/L) sum_atoms exp(2 pi i k.(r - r_a)/L)
// with Vt the 3D Fourier transform of vatom's parameterisation,
//   Vt(q) = (al/pi) sum_i a_i / (q^2 + b_i) + ag pi^(-3/2) sum_i c_i exp(-d_i q^2),
// (the transforms of exp(-2 pi r sqrt b)/r and exp(-pi^2 r^2/d)), which is
// 2 pi a0 e times Kirkland's fe(q).  k_gt_spectrum forms the band-limited
// spectrum, the axis FFTs of the post-processing path invert it, and
// k_gt_real takes the real part and the scale.  Phases are exact in
// fixed point: each coordinate is a 24-bit fraction of L, so
// k x / L mod 1 is an integer product mod 2^24.
//
// Reprojection (DHT.cpp:1509-1731, imodeReproj): the volume, converted to the
// phase of one plane by pi / (lambda E) zst (DHT.cpp:1514-1515, E the
// electron energy UTR.cpp:353-363 derives from lambda), is turned to each
// orientation (about Z, then Y', as the forward model turns atoms) and
// propagated plane by plane by the multislice kernels of msct.hip: per plane
// k_rp_trans samples the turned volume (trilinear, zero outside) and forms
// the transmission exp(i phase), then the band limit, the product with the
// wave and the propagation by zst run as for atoms.  The exit wave is
// propagated to the row's defocus distance (restated Fresnel with aperture
// and Cs, as FresnelA of DHT.cpp:1574-1577 reproduces them) and k_rp_err
// forms the reference's error of DHT.cpp:1686 against the row's image I0
// (renormalised and clamped as DHT.cpp:1641-1647):
//   err = || |psi| - sqrt(I0) ||^2 / || sqrt(I0) ||^2.
// Multislice_eV itself is in the absent XArrayLibrary: the turned volume's
// interpolation (trilinear here) is restated.

namespace utr {

constexpr double kGtPi = 3.14159265358979323846;

// Band-limited spectrum on the FFT grid [kz][ky][kx] (wrapped frequencies):
// S(k) = sum_e Vt_e(q^2) sum_{a in e} exp(-2 pi i k.r_a / L) for
// q^2 = (kx/Lx)^2 + (ky/Ly)^2 + (kz/Lz)^2 < q2max, else 0.  Atoms sorted by
// element; ux, uy, uz: coordinates as 24-bit fractions of the cube sides;
// el: [ne + 1] atom offsets per element; fe: 12 parameters per element.
__global__ void __launch_bounds__(256) k_gt_spectrum(const unsigned* __restrict__ ux, const unsigned* __restrict__ uy,
                                                     const unsigned* __restrict__ uz, const int* __restrict__ el,
                                                     int ne, const double* __restrict__ fe, int nx, int ny, int nz,
                                                     double ilx2, double ily2, double ilz2, double q2max,
                                                     cplx<float>* __restrict__ S) {
    __shared__ unsigned sx[256], sy[256], sz[256];
    const size_t n = (size_t)nx * ny * nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int kx = 0, ky = 0, kz = 0;
    bool live = false;
    double q2 = 0.0;
    if (t < n) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
        kx = i < (nx + 1) / 2 ? i : i - nx;
        ky = j < (ny + 1) / 2 ? j : j - ny;
        kz = k < (nz + 1) / 2 ? k : k - nz;
        q2 = (double)kx * kx * ilx2 + (double)ky * ky * ily2 + (double)kz * kz * ilz2;
        live = q2 < q2max;
    }
    // the whole work-group skips the atoms when none of its frequencies is kept
    if (!__syncthreads_or(live)) {
        if (t < n) S[t] = {0.f, 0.f};
        return;
    }
    const unsigned kxu = (unsigned)kx, kyu = (unsigned)ky, kzu = (unsigned)kz;
    double re = 0.0, im = 0.0;
    for (int e = 0; e < ne; e++) {
        float sre = 0.f, sim = 0.f;
        for (int a0 = el[e]; a0 < el[e + 1]; a0 += 256) {
            const int na = min(256, el[e + 1] - a0);
            __syncthreads();
            if ((int)threadIdx.x < na) {
                sx[threadIdx.x] = ux[a0 + threadIdx.x];
                sy[threadIdx.x] = uy[a0 + threadIdx.x];
                sz[threadIdx.x] = uz[a0 + threadIdx.x];
            }
            __syncthreads();
            if (live)
                for (int a = 0; a < na; a++) {
                    const unsigned ph = (kxu * sx[a] + kyu * sy[a] + kzu * sz[a]) & 0xFFFFFFu;
                    float s, c;
                    // exp(-2 pi i ph / 2^24): sincospi of 2 ph / 2^24, exact argument
                    sincospif((float)ph * (1.0f / 8388608.0f), &s, &c);
                    sre += c;
                    sim -= s;
                }
        }
        if (live) {
            const double* f = fe + 12 * e;
            double v = 0.0;
            for (int i = 0; i < 6; i += 2) v += 150.4121417 / kGtPi * f[i] / (q2 + f[i + 1]);
            for (int i = 6; i < 12; i += 2) v += 266.5985798 * pow(kGtPi, -1.5) * f[i] * exp(-f[i + 1] * q2);
            re += v * sre;
            im += v * sim;
        }
    }
    if (t < n) S[t] = {(float)re, (float)im};
}

__global__ void k_gt_real(const cplx<float>* __restrict__ S, size_t n, float scale, float* __restrict__ vol) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) vol[t] = S[t].x * scale;
}

// One orientation's turn: voxel p of the turned frame samples the volume at
// m[0..8] p + m[9..11] (voxel units).
struct RpRot {
    float m[12];
};

// Transmission of plane s for every orientation of the batch (grid: pixel
// blocks x orientations): T = exp(i phase3 V(turned)), trilinear, zero
// outside the volume.
__global__ void k_rp_trans(const float* __restrict__ vol, int nx, int ny, int nz, const RpRot* __restrict__ rot,
                           int s, float phase3, cplx<float>* __restrict__ T) {
    const int o = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int i = t % nx, j = t / nx;
    const RpRot r = rot[o];
    const float px = (float)i, py = (float)j, pz = (float)s;
    const float x = fmaf(r.m[0], px, fmaf(r.m[1], py, fmaf(r.m[2], pz, r.m[9])));
    const float y = fmaf(r.m[3], px, fmaf(r.m[4], py, fmaf(r.m[5], pz, r.m[10])));
    const float z = fmaf(r.m[6], px, fmaf(r.m[7], py, fmaf(r.m[8], pz, r.m[11])));
    float v = 0.f;
    if (x > -1.f && y > -1.f && z > -1.f && x < (float)nx && y < (float)ny && z < (float)nz) {
        const int x0 = (int)floorf(x), y0 = (int)floorf(y), z0 = (int)floorf(z);
        const float fx = x - x0, fy = y - y0, fz = z - z0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int xi = x0 + (c & 1), yi = y0 + ((c >> 1) & 1), zi = z0 + (c >> 2);
            if (xi < 0 || yi < 0 || zi < 0 || xi >= nx || yi >= ny || zi >= nz) continue;
            const float w = ((c & 1) ? fx : 1.f - fx) * (((c >> 1) & 1) ? fy : 1.f - fy) * ((c >> 2) ? fz : 1.f - fz);
            v = fmaf(w, vol[((size_t)zi * ny + yi) * nx + xi], v);
        }
    }
    float sn, cs;
    sincosf(phase3 * v, &sn, &cs);
    T[((size_t)o * ny + j) * nx + i] = {cs, sn};
}

// Per image b of the batch: num = sum (sqrt(max(I,0)) - sqrt(max(I0/f1+f2,0)))^2,
// den = sum max(I0/f1+f2, 0) (FP64, one work-group per image, fixed order).
__global__ void __launch_bounds__(256) k_rp_err(const float* __restrict__ I, const float* __restrict__ I0,
                                                size_t plane, double f1, double f2, double* __restrict__ res) {
    __shared__ double rn[256], rd[256];
    const int b = blockIdx.x;
    const float* a = I + (size_t)b * plane;
    const float* o = I0 + (size_t)b * plane;
    double num = 0.0, den = 0.0;
    for (size_t p = threadIdx.x; p < plane; p += 256) {
        double i0 = (double)o[p] / f1 + f2;
        if (i0 < 0.0) i0 = 0.0;
        const double d = sqrt(fmax((double)a[p], 0.0)) - sqrt(i0);
        num += d * d;
        den += i0;
    }
    rn[threadIdx.x] = num;
    rd[threadIdx.x] = den;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rn[threadIdx.x] += rn[threadIdx.x + w];
            rd[threadIdx.x] += rd[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        res[2 * b] = rn[0];
        res[2 * b + 1] = rd[0];
    }
}

}  // namespace utr

namespace {

// Electron energy in eV from the wavelength in Angstrom (UTR.cpp:353-363).
double ev_from_wavelength(double wl) {
    const double hp = 6.62607004e-34, cc = 299792458.0, ee = 1.602176634e-19, m0 = 9.1093837015e-31;
    const long double mc2 = (long double)m0 * cc * cc;
    const long double chl2 = (long double)(cc * cc * hp * hp) / (long double)(wl * wl * 1.e-20);
    const long double abra = std::sqrt(mc2 * mc2 + chl2);
    return (double)((-mc2 + abra) / (long double)ee);
}

}  // namespace

extern "C" {

int utr_ground_truth_device(int nx, int ny, int nz, double xst, double qmax, double scale, const double* fe_params,
                            int natom, const int* Z, const float* x, const float* y, const float* z, float* d_vol) {
    if (!d_vol || !fe_params) return set_err(UTR_E_INVALID, "utr_ground_truth_device: null volume or fe table");
    if (int rc = post_check_edges(nx, ny, nz)) return rc;
    if (!(xst > 0) || !std::isfinite(xst) || !std::isfinite(qmax) || !std::isfinite(scale))
        return set_err(UTR_E_INVALID, "utr_ground_truth_device: the voxel size must be positive and qmax, scale finite");
    if (natom < 0 || (natom > 0 && (!Z || !x || !y || !z)))
        return set_err(UTR_E_INVALID, "utr_ground_truth_device: bad atom arrays");
    for (int k = 0; k < natom; k++) {
        if (Z[k] < 1 || Z[k] > 103) return set_err(UTR_E_INVALID, "utr_ground_truth_device: atomic numbers must be in 1..103");
        if (!std::isfinite(x[k]) || !std::isfinite(y[k]) || !std::isfinite(z[k]))
            return set_err(UTR_E_INVALID, "utr_ground_truth_device: atom coordinates must be finite");
    }
    const double Lx = nx * xst, Ly = ny * xst, Lz = nz * xst;
    // atoms sorted by element, coordinates as 24-bit fractions of the sides
    std::vector<int> order(natom);
    for (int k = 0; k < natom; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return Z[a] < Z[b]; });
    std::vector<unsigned> ux(std::max(natom, 1)), uy(std::max(natom, 1)), uz(std::max(natom, 1));
    std::vector<int> el{0};
    std::vector<double> fe;
    auto frac24 = [](double v, double L) {
        double f = v / L;
        f -= std::floor(f);
        return (unsigned)((uint64_t)std::llround(f * 16777216.0) & 0xFFFFFFu);
    };
    for (int k = 0; k < natom; k++) {
        const int a = order[k];
        if (k == 0 || Z[a] != Z[order[k - 1]]) {
            if (k > 0) el.push_back(k);
            for (int i = 0; i < 12; i++) fe.push_back(fe_params[(size_t)(Z[a] - 1) * 12 + i]);
        }
        ux[k] = frac24(x[a], Lx);
        uy[k] = frac24(y[a], Ly);
        uz[k] = frac24(z[a], Lz);
    }
    el.push_back(natom);
    const int ne = natom > 0 ? (int)fe.size() / 12 : 0;
    if (fe.empty()) fe.assign(12, 0.0);
    const size_t n = (size_t)nx * ny * nz;
    LocalBuf<unsigned> dux, duy, duz;
    LocalBuf<int> del;
    LocalBuf<double> dfe;
    LocalBuf<utr::cplx<float>> S, tx, ty, tz;
    UTR_HIP_TRY(dux.alloc(ux.size()));
    UTR_HIP_TRY(duy.alloc(uy.size()));
    UTR_HIP_TRY(duz.alloc(uz.size()));
    UTR_HIP_TRY(del.alloc(el.size()));
    UTR_HIP_TRY(dfe.alloc(fe.size()));
    UTR_HIP_TRY(S.alloc(n));
    UTR_HIP_TRY(hipMemcpy(dux.p, ux.data(), sizeof(unsigned) * ux.size(), hipMemcpyHostToDevice));
    UTR_HIP_TRY(hipMemcpy(duy.p, uy.data(), sizeof(unsigned) * uy.size(), hipMemcpyHostToDevice));
    UTR_HIP_TRY(hipMemcpy(duz.p, uz.data(), sizeof(unsigned) * uz.size(), hipMemcpyHostToDevice));
    UTR_HIP_TRY(hipMemcpy(del.p, el.data(), sizeof(int) * el.size(), hipMemcpyHostToDevice));
    UTR_HIP_TRY(hipMemcpy(dfe.p, fe.data(), sizeof(double) * fe.size(), hipMemcpyHostToDevice));
    const int ns[3] = {nx, ny, nz};
    DevBuf<utr::cplx<float>>* tws[3] = {&tx, &ty, &tz};
    for (int d = 0; d < 3; d++) {
        auto tw = make_twiddles<float>(ns[d]);
        UTR_HIP_TRY(tws[d]->alloc(tw.size()));
        UTR_HIP_TRY(hipMemcpy(tws[d]->p, tw.data(), sizeof(utr::cplx<float>) * tw.size(), hipMemcpyHostToDevice));
    }
    const double q2max = qmax > 0 ? qmax * qmax : std::numeric_limits<double>::infinity();
    hipStream_t s = nullptr;
    UTR_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    auto run = [&]() -> int {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        utr::k_gt_spectrum<<<blocks, 256, 0, s>>>(dux.p, duy.p, duz.p, del.p, ne, dfe.p, nx, ny, nz, 1.0 / (Lx * Lx),
                                                   1.0 / (Ly * Ly), 1.0 / (Lz * Lz), q2max, S.p);
        UTR_HIP_TRY(hipGetLastError());
        for (int ax = 0; ax < 3; ax++)
            if (int rc = post_fft_axis(S.p, nx, ny, nz, ax, +1, tws[ax]->p, s)) return rc;
        utr::k_gt_real<<<blocks, 256, 0, s>>>(S.p, n, (float)(scale / (Lx * Ly * Lz)), d_vol);
        UTR_HIP_TRY(hipGetLastError());
        UTR_HIP_TRY(hipStreamSynchronize(s));
        return UTR_OK;
    };
    const int rc = run();
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}

int utr_reproject_device(const utr_params* p, const float* d_vol, int nrows, const double* angZ_deg,
                         const double* angY_deg, const double* defocus, const float* d_images, float* d_out,
                         double* err) {
    if (!p || !d_vol || nrows < 1 || !angZ_deg || !angY_deg || !defocus)
        return set_err(UTR_E_INVALID, "utr_reproject_device: need parameters, a volume and nrows >= 1 rows");
    if (err && !d_images) return set_err(UTR_E_INVALID, "utr_reproject_device: errors need the images");
    if (p->modality != 0) return set_err(UTR_E_UNSUPPORTED, "utr_reproject_device: electron (TEM) modality only");
    if (!(p->xst > 0) || !(p->wl > 0) || !std::isfinite(p->wl) || p->norm_f1 == 0)
        return set_err(UTR_E_INVALID, "utr_reproject_device: pixel size, wavelength and f1 must be valid");
    const int nx = p->nx, ny = p->ny;
    const int nz = (int)((p->zhi - p->zlo) / p->xst + 0.5);
    if (nx != ny || nz != nx)
        return set_err(UTR_E_UNSUPPORTED, "utr_reproject_device: the volume must be a cube (nx = ny = nz)");
    for (int r = 0; r < nrows; r++)
        if (!std::isfinite(angZ_deg[r]) || !std::isfinite(angY_deg[r]) || !std::isfinite(defocus[r]))
            return set_err(UTR_E_INVALID, "utr_reproject_device: rows must be finite");
    const double E = ev_from_wavelength(p->wl);
    const double kev = E / 1000.0;
    msct_params mp{};
    mp.nx = nx;
    mp.ny = ny;
    mp.mode = 0;
    mp.output = 0;
    mp.ctblength = nx * p->xst;
    mp.deltaz = p->xst;
    mp.kev = kev;
    mp.aobj_mrad = p->aobj_mrad;
    mp.cs3_mm = p->cs3_mm;
    mp.cs5_mm = p->cs5_mm;
    mp.dxc = p->dxc;
    mp.dyc = p->dyc;
    mp.dzc = p->dzc;
    std::vector<double> fe0(103 * 12, 0.0);
    msct_ctx* c = nullptr;
    if (int rc = msct_create(&mp, fe0.data(), 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &c)) return rc;
    // one slice per volume plane
    c->nsl = nz;
    c->g.nsl1 = nz + 1;
    const size_t plane = (size_t)nx * ny;
    const float phase3 = (float)(kMsPi / (p->wl * E) * p->xst);  // DHT.cpp:1514
    // rows grouped by defocus distance (one Fresnel table per group)
    std::vector<int> ord(nrows);
    for (int r = 0; r < nrows; r++) ord[r] = r;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return defocus[a] < defocus[b]; });
    const int B = std::min(nrows, std::max(1, c->batch));
    LocalBuf<utr::RpRot> rot;
    LocalBuf<float> img, ref;
    LocalBuf<double> res;
    hipStream_t st = nullptr;
    auto run = [&]() -> int {
        UTR_HIP_TRY(c->W.alloc(plane * B));
        UTR_HIP_TRY(c->T.alloc(plane * B));
        UTR_HIP_TRY(c->E.alloc(plane * B));
        UTR_HIP_TRY(c->F.alloc(plane));
        UTR_HIP_TRY(c->off.alloc((size_t)B * (nz + 1)));
        UTR_HIP_TRY(rot.alloc(B));
        UTR_HIP_TRY(img.alloc(plane * B));
        UTR_HIP_TRY(ref.alloc(plane * B));
        UTR_HIP_TRY(res.alloc(2 * (size_t)B));
        UTR_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        std::vector<int> off((size_t)B * (nz + 1));
        for (int o = 0; o < B; o++)
            for (int s = 0; s <= nz; s++) off[(size_t)o * (nz + 1) + s] = s;  // every plane transmits
        UTR_HIP_TRY(hipMemcpyAsync(c->off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, st));
        const double k2maxo = c->aobj > 0.f ? (double)((c->aobj / c->wavlen) * (c->aobj / c->wavlen)) : 0.0;
        const double cxv = nx / 2.0 + p->dxc / p->xst, cyv = ny / 2.0 + p->dyc / p->xst, czv = nz / 2.0 + p->dzc / p->xst;
        std::vector<utr::RpRot> R(B);
        std::vector<double> hres(2 * (size_t)B);
        double cur_def = std::numeric_limits<double>::quiet_NaN();
        for (int k0 = 0; k0 < nrows;) {
            // a batch of rows with one defocus distance
            int k1 = k0;
            while (k1 < nrows && k1 - k0 < B && defocus[ord[k1]] == defocus[ord[k0]]) k1++;
            const int nb = k1 - k0;
            if (!(defocus[ord[k0]] == cur_def)) {
                cur_def = defocus[ord[k0]];
                k_ms_fresnel<<<(unsigned)((plane + 255) / 256), 256, 0, st>>>(
                    c->F.p, nx, ny, (double)c->ctbf, (double)c->wavlen, cur_def - 0.5 * mp.ctblength,
                    (double)(float)p->cs3_mm * 1.e7, (double)(float)p->cs5_mm * 1.e7, 0.0, 1.0, 0.0, k2maxo);
                UTR_HIP_TRY(hipGetLastError());
            }
            for (int b = 0; b < nb; b++) {
                const int r = ord[k0 + b];
                const double aZ = angZ_deg[r] * kMsPi / 180.0, aY = angY_deg[r] * kMsPi / 180.0;
                const double cZ = std::cos(aZ), sZ = std::sin(aZ), cY = std::cos(aY), sY = std::sin(aY);
                // turned frame p' = c + Ry Rz (p - c)  =>  p = c + (Ry Rz)^T (p' - c)
                const double M[9] = {cY * cZ, sZ, -sY * cZ, -cY * sZ, cZ, sY * sZ, sY, 0.0, cY};
                const double cv[3] = {cxv, cyv, czv};
                for (int i = 0; i < 9; i++) R[b].m[i] = (float)M[i];
                for (int i = 0; i < 3; i++)
                    R[b].m[9 + i] = (float)(cv[i] - (M[3 * i] * cv[0] + M[3 * i + 1] * cv[1] + M[3 * i + 2] * cv[2]));
            }
            UTR_HIP_TRY(hipStreamSynchronize(st));  // the previous batch no longer reads the table
            UTR_HIP_TRY(hipMemcpyAsync(rot.p, R.data(), sizeof(utr::RpRot) * nb, hipMemcpyHostToDevice, st));
            utr::k_ms_init<<<(unsigned)((plane * nb + 255) / 256), 256, 0, st>>>(c->W.p, nx, plane * nb, 0);
            UTR_HIP_TRY(hipGetLastError());
            const float inv = 1.f / (float)plane;
            for (int s = 0; s < nz; s++) {
                utr::k_rp_trans<<<dim3((unsigned)((plane + 255) / 256), nb), 256, 0, st>>>(d_vol, nx, ny, nz, rot.p, s,
                                                                                         phase3, c->T.p);
                UTR_HIP_TRY(hipGetLastError());
                int rc;
                if ((rc = ms_run_lines(c, 0, nb, s, 0, 0, nullptr, st))) return rc;
                if ((rc = ms_run_cols(0, c, c->T.p, c->T.p, nb, inv, s, nullptr, st))) return rc;
                if ((rc = ms_run_lines(c, 1, nb, s, 0, 0, nullptr, st))) return rc;
                if ((rc = ms_run_cols(1, c, c->W.p, c->W.p, nb, inv, s, nullptr, st))) return rc;
            }
            if (int rc = ms_run_cols(2, c, c->W.p, c->E.p, nb, 1.f / (float)ny, 0, c->F.p, st)) return rc;
            if (int rc = ms_run_lines(c, 3, nb, 0, 1, 0, img.p, st)) return rc;
            for (int b = 0; b < nb; b++) {
                const int r = ord[k0 + b];
                if (d_out)
                    UTR_HIP_TRY(hipMemcpyAsync(d_out + plane * r, img.p + plane * b, sizeof(float) * plane,
                                               hipMemcpyDeviceToDevice, st));
                if (err)
                    UTR_HIP_TRY(hipMemcpyAsync(ref.p + plane * b, d_images + plane * r, sizeof(float) * plane,
                                               hipMemcpyDeviceToDevice, st));
            }
            if (err) {
                utr::k_rp_err<<<nb, 256, 0, st>>>(img.p, ref.p, plane, p->norm_f1, p->norm_f2, res.p);
                UTR_HIP_TRY(hipGetLastError());
                UTR_HIP_TRY(hipMemcpyAsync(hres.data(), res.p, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st));
                UTR_HIP_TRY(hipStreamSynchronize(st));
                for (int b = 0; b < nb; b++) err[ord[k0 + b]] = hres[2 * b + 1] > 0 ? hres[2 * b] / hres[2 * b + 1] : 0.0;
            }
            k0 = k1;
        }
        UTR_HIP_TRY(hipStreamSynchronize(st));
        return UTR_OK;
    };
    const int rc = run();
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    msct_destroy(c);
    return rc;
}

}  // extern "C"
