// utr_api.hip -- C ABI of libutr_hip.so: context, device buffers, launches.
//
// Replaces the FFTW plans and the OpenMP orientation loop of
// MultisliceCppTest/PhaseRetrieval/UTR.cpp:840-1279 (see include/utr_hip.h).
// Three forms share the FFT and epilogue kernels:
//   UTR_FP32         rows are transformed in batches as they arrive; finalize
//                    runs the brick scatter (exact fixed-point sums in LDS,
//                    deterministic), the fused normalisation/filters and the
//                    3D c2r.
//   UTR_FP32_GATHER  as UTR_FP32 with a per-voxel gather and Kahan-compensated
//                    FP32 sums instead of the brick scatter.
//   UTR_FP64         the plain form: one row at a time, scatter gridding with
//                    FP64 atomics, then separate normalisation, symmetrisation
//                    and c2r passes (the structure of the reference's loop, in
//                    double).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <limits>
#include <vector>
#include <atomic>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>

#include "../../include/utr_hip.h"
#include "utr_fft.hpp"
#include "utr_kernels.hip"
#include "utr_plan.hpp"

using utr::cplx;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define UTR_HIP_TRY(expr)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err(e_ == hipErrorOutOfMemory ? UTR_E_NOMEM : UTR_E_DEVICE,             \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

bool fft_size_supported(int n) {
    switch (n) {
        case 16: case 32: case 48: case 64: case 96: case 128: case 160: case 192:
        case 256: case 320: case 384: case 512:
            return true;
        default:
            return false;
    }
}

// Dispatches F<N>() for the supported FFT lengths.
#define UTR_DISPATCH_N(n, F)                  \
    switch (n) {                              \
        case 16: F(16); break;                \
        case 32: F(32); break;                \
        case 48: F(48); break;                \
        case 64: F(64); break;                \
        case 96: F(96); break;                \
        case 128: F(128); break;              \
        case 160: F(160); break;              \
        case 192: F(192); break;              \
        case 256: F(256); break;              \
        case 320: F(320); break;              \
        case 384: F(384); break;              \
        case 512: F(512); break;              \
        default: break;                       \
    }

template <typename T>
std::vector<cplx<T>> make_twiddles(int n) {
    const utr::FacList F = utr::factorize(n);
    std::vector<cplx<T>> tw(F.ntw > 0 ? F.ntw : 1);
    for (int s = 0; s < F.nf; s++) {
        const int p = F.p[s], R = F.f[s];
        for (int k = 0; k < p; k++)
            for (int r = 1; r < R; r++) {
                const double ang = -2.0 * utr::kPi * (double)r * k / ((double)p * R);
                tw[F.off[s] + k * (R - 1) + (r - 1)] = {T(std::cos(ang)), T(std::sin(ang))};
            }
    }
    return tw;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        if (p && n >= count) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// DevBuf owned by one scope (function-local scratch): freed on every return
// path, error returns included.
template <typename T>
struct LocalBuf : DevBuf<T> {
    LocalBuf() = default;
    LocalBuf(const LocalBuf&) = delete;
    LocalBuf& operator=(const LocalBuf&) = delete;
    ~LocalBuf() { this->release(); }
};

constexpr int kStageRows = 32;  // rows per host->device staging batch
constexpr int kRowsWgPerCu = 8; // resident workgroups of the persistent row FFT

// ------------------------------------------------ host memory, host copies
//
// Pinned host memory handed out by utr_host_alloc, by address range: the
// host-memory entry points move data between such memory and the device by
// DMA directly.  Other (pageable) host memory goes through pinned staging
// slots owned by the context, filled and drained by HostCopier.

std::mutex g_pinned_mu;
std::map<uintptr_t, size_t> g_pinned;  // base -> bytes

bool host_pinned(const void* p, size_t bytes) {
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    auto it = g_pinned.upper_bound(a);
    if (it == g_pinned.begin()) return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}

// A few host threads copying one buffer together: one thread's memcpy from
// pageable memory reaches a few GB/s, well below the DMA rate of the pinned
// slots it feeds.  The calling thread takes part; copy() returns when the
// whole buffer is copied.  One process-wide instance (calls serialise).
class HostCopier {
public:
    static HostCopier& get() {
        static HostCopier inst(threads());
        return inst;
    }
    void copy(void* dst, const void* src, size_t bytes) {
        std::lock_guard<std::mutex> one(call_mu_);
        if (bytes < (kPiece << 1) || th_.empty()) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = (char*)dst;
            src_ = (const char*)src;
            bytes_ = bytes;
            next_.store(0);
            busy_ = (int)th_.size();
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return busy_ == 0; });
    }
    ~HostCopier() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }

private:
    static constexpr size_t kPiece = size_t(1) << 20;
    static int threads() {
        // UTR_HOST_THREADS overrides; the caller is one of them
        const char* e = std::getenv("UTR_HOST_THREADS");
        int n = e && *e ? std::atoi(e) : (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
        return std::max(1, std::min(n, 64));
    }
    explicit HostCopier(int n) {
        for (int i = 1; i < n; i++) th_.emplace_back([this] { run(); });
    }
    void work() {
        for (;;) {
            const size_t o = next_.fetch_add(kPiece);
            if (o >= bytes_) break;
            std::memcpy(dst_ + o, src_ + o, std::min(kPiece, bytes_ - o));
        }
    }
    void run() {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> lk(m_);
            if (--busy_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_, call_mu_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    std::atomic<size_t> next_{0};
    unsigned gen_ = 0;
    int busy_ = 0;
    bool stop_ = false;
};

// Three pinned staging slots of `slot` floats with an event per slot (the
// slot is free once the copy recorded after its last use has completed).
struct PinnedStage {
    float* p = nullptr;
    size_t slot = 0;
    hipEvent_t ev[3] = {};
    hipError_t ensure(size_t floats, hipStream_t s) {
        if (p && slot >= floats) return hipSuccess;
        release();
        hipError_t e = hipHostMalloc((void**)&p, sizeof(float) * floats * 3, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        slot = floats;
        for (auto& v : ev) {
            if ((e = hipEventCreateWithFlags(&v, hipEventDisableTiming)) != hipSuccess) return e;
            if ((e = hipEventRecord(v, s)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    float* at(int i) const { return p + (size_t)i * slot; }
    void release() {
        for (auto& v : ev)
            if (v) (void)hipEventDestroy(v);
        for (auto& v : ev) v = nullptr;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        slot = 0;
    }
};

}  // namespace

struct utr_ctx {
    utr::Plan plan;
    int precision = UTR_FP32;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    int ncu = 256;  // compute units of the device (persistent grids)
    std::vector<char> added;
    int n_added = 0;
    int keep_sampling = 0;
    bool finalized = false;
    utr_timings tim = {};
    float add_ms = 0;  // accumulated device time of add_images (2D stage + FP64 scatter)
    // shared
    DevBuf<utr::RowParams> rows;
    DevBuf<float> img_stage;
    PinnedStage hstage;  // pinned slots for pageable host images and volumes
    // input image correction (Z'' rotation, RELION shift) of the rows that need it
    DevBuf<utr::PrepRow> prep;
    DevBuf<float> img_prep, prep_aver;
    // FP32 form
    DevBuf<cplx<float>> spec32, V32, H32, twx32, twy32, twz32;
    DevBuf<float> S32, vol32;
    DevBuf<unsigned> rowmax;  // per-row max |K| as float bits (FP32 forms)
    int grid_planes = 0;      // y planes per workgroup of the plane gridding (0: by size)
    int rows_wpc = 0;         // resident workgroups per CU of the persistent row FFT (0: default)
    int cols_wpc = 0;         // same for the column FFT (0: as many as LDS allows)
    DevBuf<int> tile_order;   // plane-gridding tiles by decreasing expected work
    DevBuf<double> rowc;      // per row: m0 m1 m2 m6 m8 (padded to 6) for the plane gridding
    // FP64 form
    DevBuf<cplx<double>> spec64, V64, H64, twx64, twy64, twz64;
    DevBuf<double> S64, vol64;
};

namespace {

size_t vox(const utr_ctx* c) {
    const auto& g = c->plan.g;
    return (size_t)g.nx * g.ny * g.nz;
}

// Grid of a persistent column-type kernel: wpc resident workgroups per CU,
// rounded down to a multiple of 8 (the XCD count) as its tile order needs.
unsigned persistent_cap(const utr_ctx* c, int wpc) {
    const unsigned n = (unsigned)(c->ncu * wpc) & ~7u;
    return n ? n : 8u;
}

template <typename T>
int upload_twiddles(utr_ctx* c, DevBuf<cplx<T>>& bx, DevBuf<cplx<T>>& by, DevBuf<cplx<T>>& bz) {
    const auto& g = c->plan.g;
    const int ns[3] = {g.nx, g.ny, g.nz};
    DevBuf<cplx<T>>* bs[3] = {&bx, &by, &bz};
    for (int d = 0; d < 3; d++) {
        auto tw = make_twiddles<T>(ns[d]);
        UTR_HIP_TRY(bs[d]->alloc(tw.size()));
        UTR_HIP_TRY(hipMemcpy(bs[d]->p, tw.data(), sizeof(cplx<T>) * tw.size(), hipMemcpyHostToDevice));
    }
    return UTR_OK;
}

// 2D stage for `count` rows whose images are at d_img (device).
template <typename T>
int launch_2d(utr_ctx* c, const float* d_img, int row0, int count, cplx<T>* d_spec,
              const cplx<T>* twx, const cplx<T>* twy) {
    const auto& g = c->plan.g;
    if (count <= 0) return UTR_OK;
    bool ok = false;
#define UTR_ROWS(N)                                                                             \
    {                                                                                           \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<T>), utr::LDS_ROWS);                      \
        const int nbr = (g.ny + NB - 1) / NB, nbt = nbr * count;                                \
        const int grid = std::min(nbt, c->ncu * (c->rows_wpc > 0 ? c->rows_wpc : kRowsWgPerCu)); \
        utr::k_rows<T, N><<<grid, 256, 0, c->stream>>>(d_img, d_spec, g.ny, g, twx, nbr, nbt);  \
        ok = true;                                                                              \
    }
    UTR_DISPATCH_N(g.nx, UTR_ROWS)
#undef UTR_ROWS
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image width for the FFT");
    ok = false;
#define UTR_COLS(N)                                                                             \
    {                                                                                           \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<T>), utr::LDS_COLS);                      \
        const int ncol = g.nx / 2 + 1; /* the rest by Hermitian symmetry */                     \
        const int gx = (ncol + NB - 1) / NB;                                                    \
        const int lds = 2 * utr::lds_len(NB * N) * (int)sizeof(cplx<T>);                          \
        int wpc = std::max(1, std::min(8, 163840 / lds)); /* resident per CU */                 \
        if (c->cols_wpc > 0) wpc = std::min(wpc, c->cols_wpc);                                  \
        const unsigned grid = std::min(utr::xcd_grid(gx * count), persistent_cap(c, wpc));      \
        utr::k_cols<T, N><<<grid, 256, 0, c->stream>>>(d_spec, g.nx, ncol, g, c->rows.p, row0,  \
                                                       twy, c->rowmax.p, gx, count);            \
        ok = true;                                                                              \
    }
    UTR_DISPATCH_N(g.ny, UTR_COLS)
#undef UTR_COLS
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image height for the FFT");
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// 3D inverse: pack + z, y, x c2r with mask and scaling.
template <typename T, typename TO>
int launch_3d(utr_ctx* c, const cplx<T>* d_V, cplx<T>* d_H, TO* d_vol, const cplx<T>* twx,
              const cplx<T>* twy, const cplx<T>* twz) {
    const auto& g = c->plan.g;
    const int W = g.nx / 2 + 1;
    bool ok = false;
#define UTR_PZ(N)                                                                               \
    {                                                                                           \
        constexpr int CW = utr::lds_batch(N, sizeof(cplx<T>), utr::LDS_PACKZ);                     \
        const int gx = (W + CW - 1) / CW;                                                       \
        utr::k_pack_z<T, N><<<utr::xcd_grid(gx * g.ny), 256, 0, c->stream>>>(d_V, d_H, g, twz,  \
                                                                             gx, g.ny);         \
        ok = true;                                                                              \
    }
    UTR_DISPATCH_N(g.nz, UTR_PZ)
#undef UTR_PZ
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported volume depth for the FFT");
    ok = false;
#define UTR_FY(N)                                                                               \
    {                                                                                           \
        constexpr int CW = utr::lds_batch(N, sizeof(cplx<T>), utr::LDS_FFTY);                      \
        const int gx = (W + CW - 1) / CW;                                                       \
        utr::k_fft_y<T, N><<<utr::xcd_grid(gx * g.nz), 256, 0, c->stream>>>(d_H, g, twy, gx,    \
                                                                            g.nz);              \
        ok = true;                                                                              \
    }
    UTR_DISPATCH_N(g.ny, UTR_FY)
#undef UTR_FY
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image height for the FFT");
    ok = false;
#define UTR_CX(N)                                                                               \
    {                                                                                           \
        constexpr int NB = utr::lds_batch(N, sizeof(cplx<T>), utr::LDS_C2R);                       \
        const int gx = (g.ny / 2 + NB - 1) / NB;                                                \
        utr::k_c2r_x<T, N, TO><<<utr::xcd_grid(gx * g.nz), 256, 0, c->stream>>>(d_H, d_vol, g,  \
                                                                               twx, gx, g.nz);  \
        ok = true;                                                                              \
    }
    UTR_DISPATCH_N(g.nx, UTR_CX)
#undef UTR_CX
    if (!ok) return set_err(UTR_E_UNSUPPORTED, "unsupported image width for the FFT");
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// Launches the 2D stage (and, in the plain FP64 form, the scatter) of `count`
// rows whose images are at d_img (device), without waiting.
constexpr int kPrepRows = 64;  // images corrected per pass (img_prep holds this many)

// Device buffers of the input image correction, when some row needs it.
int alloc_prep(utr_ctx* c) {
    bool any = false;
    for (const auto& q : c->plan.prep) any |= q.on != 0;
    if (!any) return UTR_OK;
    const auto& g = c->plan.g;
    const int k = std::min(kPrepRows, c->plan.nrows);
    UTR_HIP_TRY(c->img_prep.alloc((size_t)k * g.nx * g.ny));
    UTR_HIP_TRY(c->prep_aver.alloc(k));
    return UTR_OK;
}

// Input image correction of count <= kPrepRows rows (UTR.cpp:960-977) into
// img_prep: the edge mean of each image, then its shifted and rotated copy
// (rows without a correction are copied unchanged).
int prep_images(utr_ctx* c, int first_row, int count, const float* d_img) {
    const auto& g = c->plan.g;
    const size_t n2 = (size_t)g.nx * g.ny;
    if (count > kPrepRows || c->img_prep.n < n2 * count || c->prep_aver.n < (size_t)count)
        return set_err(UTR_E_STATE, "image correction buffers not allocated");
    utr::k_prep_aver<<<count, 256, 0, c->stream>>>(d_img, c->prep.p, first_row, g, c->prep_aver.p);
    UTR_HIP_TRY(hipGetLastError());
    // rotation centre (UTR.cpp:342-343) in pixels: pixel i sits at xlo + i xst
    const double px = (c->plan.xc - g.xlo) / c->plan.xst;
    const double py = (c->plan.yc - g.ylo) / c->plan.xst;
    const dim3 grid((unsigned)((n2 + 255) / 256), (unsigned)count);
    utr::k_prep_img<<<grid, 256, 0, c->stream>>>(d_img, c->img_prep.p, c->prep.p, first_row, g, px, py,
                                                 c->prep_aver.p);
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

// stage(first, count, images) for rows [first_row, first_row + count) whose
// images are at d_img, through the input image correction where a row
// needs it.
template <typename F>
int with_prep(utr_ctx* c, int first_row, int count, const float* d_img, F&& stage) {
    bool any = false;
    for (int r = 0; r < count; r++) any |= c->plan.prep[first_row + r].on != 0;
    if (!any) return stage(first_row, count, d_img);
    const size_t n2 = (size_t)c->plan.g.nx * c->plan.g.ny;
    for (int r = 0; r < count; r += kPrepRows) {
        const int k = std::min(kPrepRows, count - r);
        int rc = prep_images(c, first_row + r, k, d_img + r * n2);
        if (rc) return rc;
        rc = stage(first_row + r, k, c->img_prep.p);
        if (rc) return rc;
    }
    return UTR_OK;
}

int enqueue_plain(utr_ctx* c, int first_row, int count, const float* d_img) {
    const auto& g = c->plan.g;
    if (c->precision != UTR_FP64) {
        return launch_2d<float>(c, d_img, first_row, count,
                                c->spec32.p + (size_t)first_row * g.ny * g.nx, c->twx32.p, c->twy32.p);
    }
    // plain form: one row at a time, 2D stage then scatter
    const size_t n2 = (size_t)g.nx * g.ny;
    for (int r = 0; r < count; r++) {
        int rc = launch_2d<double>(c, d_img + r * n2, first_row + r, 1, c->spec64.p, c->twx64.p,
                                   c->twy64.p);
        if (rc) return rc;
        const int threads = 256;
        const int blocks = (int)((n2 + threads - 1) / threads);
        utr::k_scatter<<<blocks, threads, 0, c->stream>>>(c->spec64.p, c->rows.p, first_row + r, g,
                                                          (double*)c->V64.p, c->S64.p);
        UTR_HIP_TRY(hipGetLastError());
    }
    return UTR_OK;
}

int enqueue_rows(utr_ctx* c, int first_row, int count, const float* d_img) {
    return with_prep(c, first_row, count, d_img,
                     [&](int f, int k, const float* im) { return enqueue_plain(c, f, k, im); });
}

// Waits for the rows enqueued since event e0, adds their time to the 2D
// phase and marks them added.
int finish_rows(utr_ctx* c, hipEvent_t e0, int first_row, int count) {
    hipEvent_t e1 = c->ev[7];
    UTR_HIP_TRY(hipEventRecord(e1, c->stream));
    UTR_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    UTR_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    c->add_ms += ms;
    for (int r = 0; r < count; r++) {
        if (!c->added[first_row + r]) c->n_added++;
        c->added[first_row + r] = 1;
    }
    return UTR_OK;
}

int add_rows(utr_ctx* c, int first_row, int count, const float* d_img) {
    hipEvent_t e0 = c->ev[6];
    UTR_HIP_TRY(hipEventRecord(e0, c->stream));
    const int rc = enqueue_rows(c, first_row, count, d_img);
    if (rc) return rc;
    return finish_rows(c, e0, first_row, count);
}

// Device -> host copy of n floats into the caller's memory: one DMA into
// pinned caller memory; otherwise in pieces through the pinned slots, each
// piece drained by the host threads while the DMA of the next two runs.
int copy_to_host(utr_ctx* c, float* dst, const float* d_src, size_t n) {
    if (host_pinned(dst, sizeof(float) * n)) {
        UTR_HIP_TRY(hipMemcpyAsync(dst, d_src, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        UTR_HIP_TRY(hipStreamSynchronize(c->stream));
        return UTR_OK;
    }
    const size_t piece = size_t(2) << 20;  // floats (8 MB)
    UTR_HIP_TRY(c->hstage.ensure(std::max<size_t>(piece, c->hstage.slot), c->stream));
    const size_t np = (n + piece - 1) / piece;
    auto issue = [&](size_t i) -> hipError_t {
        const size_t off = i * piece, m = std::min(piece, n - off);
        hipError_t e = hipMemcpyAsync(c->hstage.at((int)(i % 3)), d_src + off, sizeof(float) * m,
                                      hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->hstage.ev[i % 3], c->stream);
        return e;
    };
    for (size_t i = 0; i < np && i < 2; i++) UTR_HIP_TRY(issue(i));
    for (size_t i = 0; i < np; i++) {
        if (i + 2 < np) UTR_HIP_TRY(issue(i + 2));
        UTR_HIP_TRY(hipEventSynchronize(c->hstage.ev[i % 3]));
        const size_t off = i * piece, m = std::min(piece, n - off);
        HostCopier::get().copy(dst + off, c->hstage.at((int)(i % 3)), sizeof(float) * m);
    }
    return UTR_OK;
}

// The plane gridding's tiles by increasing distance of their centre from
// the plane's rotation centre (nx/2, nz/2): the point density of the lines
// falls off as 1/r, so the nearest tiles carry the most points and start
// first.  Uploaded once at creation (finalize may run under graph capture).
int upload_plane_tables(utr_ctx* c) {
    const auto& g = c->plan.g;
    const int ntx = (g.nx + utr::PX - 1) / utr::PX, ntz = (g.nz + utr::PZ - 1) / utr::PZ;
    const int ntiles = ntx * ntz;
    std::vector<int> ord(ntiles);
    std::vector<double> d2(ntiles);
    for (int t = 0; t < ntiles; t++) {
        ord[t] = t;
        const double cx = (t % ntx) * utr::PX + 0.5 * utr::PX - g.nxd2;
        const double cz = (t / ntx) * utr::PZ + 0.5 * utr::PZ - g.nzd2;
        d2[t] = cx * cx + cz * cz;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d2[a] < d2[b]; });
    UTR_HIP_TRY(c->tile_order.alloc(ntiles));
    UTR_HIP_TRY(hipMemcpy(c->tile_order.p, ord.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice));
    // the map coefficients the points read, 48 bytes a row
    const int nr = c->plan.nrows;
    std::vector<double> rc(6 * (size_t)nr, 0.0);
    for (int r = 0; r < nr; r++) {
        const double* m = c->plan.rows[r].m;
        rc[6 * r + 0] = m[0]; rc[6 * r + 1] = m[1]; rc[6 * r + 2] = m[2];
        rc[6 * r + 3] = m[6]; rc[6 * r + 4] = m[8];
    }
    UTR_HIP_TRY(c->rowc.alloc(rc.size()));
    UTR_HIP_TRY(hipMemcpy(c->rowc.p, rc.data(), sizeof(double) * rc.size(), hipMemcpyHostToDevice));
    return UTR_OK;
}

int check_rows(utr_ctx* c, int first_row, int count) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (first_row < 0 || count < 0 || first_row + count > c->plan.nrows)
        return set_err(UTR_E_INVALID, "row range outside the orientation table");
    return UTR_OK;
}

// Gridding of the FP32 form on the context's stream: y-plane tiles when no
// row rotates about Z, bricks otherwise.  UTR_GRID_FORM=brick in the
// environment selects the bricks for every job (both give the same sums).
// ev_start (may be null): recorded on the stream right before the gridding
// kernel itself, so that the phase timer covers that kernel alone.
int launch_grid32(utr_ctx* c, float* d_S, hipEvent_t ev_start) {
    const auto& g = c->plan.g;
    const char* form = std::getenv("UTR_GRID_FORM");
    const bool brick = form && std::strcmp(form, "brick") == 0;
    if (c->plan.all_y_exact && !brick) {
        const long ntiles = (long)((g.nx + utr::PX - 1) / utr::PX) * ((g.nz + utr::PZ - 1) / utr::PZ);
        int pg = c->grid_planes;
        if (pg <= 0) {
            const char* e = std::getenv("UTR_GRID_PLANES");
            pg = e ? std::atoi(e) : 0;
        }
        // default: groups of up to 4 planes while the launch keeps at least
        // four workgroups per resident slot (4 per CU)
        if (pg <= 0) pg = (int)std::max(1L, std::min(4L, ntiles * g.ny / (16L * c->ncu)));
        const long ngroups = (g.ny + pg - 1) / pg;
        const long nwg = 8L * ((ngroups + 7) / 8) * ntiles;
        if (nwg <= 0 || nwg > 0x7ffffff0L) return set_err(UTR_E_INVALID, "volume too large");
        if (c->tile_order.n != (size_t)ntiles || c->rowc.n != 6 * (size_t)c->plan.nrows)
            return set_err(UTR_E_STATE, "plane-gridding tables missing");
        // the noise gate alone in the epilogue (the metric's mode 0 with NSR)
        const bool gate = g.nsr2 != 0 && g.ctf_mode != 1 && !g.dht_sctf && !g.do_tie3d &&
                          !g.do_invlap && !g.do_ctf3d;
        const unsigned nb = utr::yp_grid((int)ntiles, (int)ngroups);
        // plane groups per block of an XCD's launch order (all: heaviest tiles
        // of all its groups first)
        int gblk = 0;
        {
            const char* e = std::getenv("UTR_GRID_GBLK");
            gblk = e ? std::atoi(e) : 0;
        }
        const int ngx = (int)((ngroups + 7) / 8);
        if (gblk <= 0 || gblk > ngx) gblk = ngx;
        if (ev_start) UTR_HIP_TRY(hipEventRecord(ev_start, c->stream));
#define UTR_GRID_YP(ROT, GATE, PL)                                                               \
    utr::k_grid_yp<ROT, GATE, PL><<<nb, 256, 0, c->stream>>>(c->spec32.p, c->rows.p, c->rowc.p,  \
                                                             c->plan.nrows, c->rowmax.p, g,       \
                                                             c->V32.p, d_S, (int)ntiles,          \
                                                             (int)ngroups, pg, c->tile_order.p,   \
                                                             gblk)
        // list capacity: one list per plane group up to 768 rows
        const bool small = c->plan.nrows <= utr::P_LIST_S;
        if (g.rot_centre_shift) {
            if (small) UTR_GRID_YP(true, false, utr::P_LIST_S);
            else UTR_GRID_YP(true, false, utr::P_LIST_L);
        } else if (gate) {
            if (small) UTR_GRID_YP(false, true, utr::P_LIST_S);
            else UTR_GRID_YP(false, true, utr::P_LIST_L);
        } else {
            if (small) UTR_GRID_YP(false, false, utr::P_LIST_S);
            else UTR_GRID_YP(false, false, utr::P_LIST_L);
        }
#undef UTR_GRID_YP
    } else {
        const long nblocks = (long)((g.nx + utr::SBX - 1) / utr::SBX) *
                             ((g.ny + utr::SBY - 1) / utr::SBY) * ((g.nz + utr::SBZ - 1) / utr::SBZ);
        if (nblocks <= 0 || nblocks > 0x7fffffffL) return set_err(UTR_E_INVALID, "volume too large");
        if (ev_start) UTR_HIP_TRY(hipEventRecord(ev_start, c->stream));
        utr::k_grid_fx<<<(unsigned)nblocks, 256, 0, c->stream>>>(c->spec32.p, c->rows.p, c->plan.nrows,
                                                                 c->rowmax.p, g, c->V32.p, d_S);
    }
    UTR_HIP_TRY(hipGetLastError());
    return UTR_OK;
}

int do_finalize(utr_ctx* c, float* d_vol_out, double* d_vol64_out) {
    const auto& g = c->plan.g;
    if (c->n_added != c->plan.nrows)
        return set_err(UTR_E_STATE, "utr_finalize: " + std::to_string(c->plan.nrows - c->n_added) +
                                        " orientation rows have no image");
    const size_t N = vox(c);
    if (c->precision != UTR_FP32) UTR_HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    if (c->precision != UTR_FP64) {
        if (c->keep_sampling) UTR_HIP_TRY(c->S32.alloc(N));
        float* d_S = c->keep_sampling ? c->S32.p : nullptr;
        if (c->precision == UTR_FP32) {
            const int rc = launch_grid32(c, d_S, c->ev[0]);
            if (rc) return rc;
        } else {
            const long nblocks = (long)((g.nx + utr::GBX - 1) / utr::GBX) *
                                 ((g.ny + utr::GBY - 1) / utr::GBY) * ((g.nz + utr::GBZ - 1) / utr::GBZ);
            if (nblocks <= 0 || nblocks > 0x7fffffffL) return set_err(UTR_E_INVALID, "volume too large");
            utr::k_gather<<<(unsigned)nblocks, 256, 0, c->stream>>>(c->spec32.p, c->rows.p,
                                                                   c->plan.nrows, g, c->V32.p, d_S);
        }
        UTR_HIP_TRY(hipGetLastError());
        UTR_HIP_TRY(hipEventRecord(c->ev[1], c->stream));
        int rc = launch_3d<float, float>(c, c->V32.p, c->H32.p, d_vol_out, c->twx32.p, c->twy32.p,
                                         c->twz32.p);
        if (rc) return rc;
    } else {
        const int threads = 256;
        const unsigned blocks = (unsigned)((N + threads - 1) / threads);
        utr::k_normalize<double><<<blocks, threads, 0, c->stream>>>(c->V64.p, c->S64.p, g);
        UTR_HIP_TRY(hipGetLastError());
        UTR_HIP_TRY(hipEventRecord(c->ev[1], c->stream));
        int rc;
        if (d_vol64_out)
            rc = launch_3d<double, double>(c, c->V64.p, c->H64.p, d_vol64_out, c->twx64.p,
                                           c->twy64.p, c->twz64.p);
        else
            rc = launch_3d<double, float>(c, c->V64.p, c->H64.p, d_vol_out, c->twx64.p,
                                          c->twy64.p, c->twz64.p);
        if (rc) return rc;
    }
    UTR_HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    UTR_HIP_TRY(hipEventSynchronize(c->ev[2]));
    float a = 0, b = 0;
    UTR_HIP_TRY(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    UTR_HIP_TRY(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    c->tim.backprop_ms = a;
    c->tim.fft3d_ms = b;
    c->tim.fft2d_ctf_ms = c->add_ms;
    c->tim.total_ms = c->add_ms + a + b;
    c->finalized = true;
    return UTR_OK;
}

}  // namespace

extern "C" {

const char* utr_last_error(void) { return g_err.c_str(); }

int utr_abi_version(void) { return UTR_ABI_VERSION; }

int utr_angle_coeffs(const utr_params* params, double angZ_deg, double angY_deg, double* m9) {
    const double z = 0.0, zero = 0.0;
    utr::Plan P;
    std::string err;
    if (!utr::build_plan(params, 1, &z, &zero, nullptr, &zero, nullptr, P, err))
        return set_err(UTR_E_INVALID, err);
    utr::angle_coeffs(P.g, angZ_deg, angY_deg, m9);
    return UTR_OK;
}

int utr_create(const utr_params* params, int nrows, const double* angZ_deg, const double* angY_deg,
               const double* angZ2_deg, const double* defocus, const float* weights, int precision,
               utr_ctx** out) {
    if (!out) return set_err(UTR_E_INVALID, "null output pointer");
    *out = nullptr;
    if (precision != UTR_FP32 && precision != UTR_FP64 && precision != UTR_FP32_GATHER)
        return set_err(UTR_E_INVALID, "precision must be UTR_FP32, UTR_FP64 or UTR_FP32_GATHER");
    utr_ctx* c = new utr_ctx();
    std::string err;
    if (!utr::build_plan(params, nrows, angZ_deg, angY_deg, angZ2_deg, defocus, weights, c->plan, err)) {
        delete c;
        return set_err(UTR_E_INVALID, err);
    }
    const auto& g = c->plan.g;
    if (!fft_size_supported(g.nx) || !fft_size_supported(g.ny) || !fft_size_supported(g.nz)) {
        delete c;
        return set_err(UTR_E_UNSUPPORTED,
                       "volume edge not supported by the FFT kernels (supported: 16 32 48 64 96 128 "
                       "160 192 256 320 384 512)");
    }
    c->precision = precision;
    c->added.assign(nrows, 0);
    auto fail = [&](int rc) {
        utr_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return fail(set_err(UTR_E_DEVICE, "no HIP device"));
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        c->ncu <= 0)
        return fail(set_err(UTR_E_DEVICE, "hipDeviceGetAttribute failed"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "hipStreamCreate failed"));
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(set_err(UTR_E_DEVICE, "hipEventCreate failed"));
    if (c->rows.alloc(nrows) != hipSuccess) return fail(set_err(UTR_E_NOMEM, "row table allocation"));
    if (hipMemcpy(c->rows.p, c->plan.rows.data(), sizeof(utr::RowParams) * nrows,
                  hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "row table upload"));
    if (c->prep.alloc(nrows) != hipSuccess ||
        hipMemcpy(c->prep.p, c->plan.prep.data(), sizeof(utr::PrepRow) * nrows, hipMemcpyHostToDevice) !=
            hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "image correction table upload"));
    if (alloc_prep(c)) return fail(set_err(UTR_E_NOMEM, "image correction buffers"));
    const size_t N = vox(c), n2 = (size_t)g.nx * g.ny;
    const size_t nh = (size_t)g.nz * g.ny * utr::h_stride(g.nx);
    int rc;
    if (precision != UTR_FP64) {
        if ((rc = upload_twiddles<float>(c, c->twx32, c->twy32, c->twz32))) return fail(rc);
        if (c->spec32.alloc((size_t)nrows * n2) != hipSuccess || c->V32.alloc(N) != hipSuccess ||
            c->H32.alloc(nh) != hipSuccess || c->rowmax.alloc(nrows) != hipSuccess)
            return fail(set_err(UTR_E_NOMEM, "device allocation of the FP32 buffers failed"));
        if (c->plan.all_y_exact) {
            const int rc = upload_plane_tables(c);
            if (rc) return fail(rc);
        }
        // on the context's stream: it is non-blocking, so a null-stream memset
        // would not be ordered before the kernels launched on it
        if (hipMemsetAsync(c->rowmax.p, 0, sizeof(unsigned) * nrows, c->stream) != hipSuccess)
            return fail(set_err(UTR_E_DEVICE, "hipMemsetAsync failed"));
    } else {
        if ((rc = upload_twiddles<double>(c, c->twx64, c->twy64, c->twz64))) return fail(rc);
        if (c->spec64.alloc(n2) != hipSuccess || c->V64.alloc(N) != hipSuccess ||
            c->S64.alloc(N) != hipSuccess || c->H64.alloc(nh) != hipSuccess)
            return fail(set_err(UTR_E_NOMEM, "device allocation of the FP64 buffers failed"));
        if (hipMemsetAsync(c->V64.p, 0, sizeof(cplx<double>) * N, c->stream) != hipSuccess ||
            hipMemsetAsync(c->S64.p, 0, sizeof(double) * N, c->stream) != hipSuccess)
            return fail(set_err(UTR_E_DEVICE, "hipMemsetAsync failed"));
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "hipStreamSynchronize failed"));
    *out = c;
    return UTR_OK;
}

int utr_add_images_device(utr_ctx* c, int first_row, int count, const float* d_images) {
    int rc = check_rows(c, first_row, count);
    if (rc) return rc;
    if (count == 0) return UTR_OK;
    if (!d_images) return set_err(UTR_E_INVALID, "null image pointer");
    return add_rows(c, first_row, count, d_images);
}

int utr_add_images(utr_ctx* c, int first_row, int count, const float* images) {
    int rc = check_rows(c, first_row, count);
    if (rc) return rc;
    if (count == 0) return UTR_OK;
    if (!images) return set_err(UTR_E_INVALID, "null image pointer");
    const auto& g = c->plan.g;
    const size_t n2 = (size_t)g.nx * g.ny, slot = n2 * kStageRows;
    // Two device staging slots: the copy of batch b+1 overlaps the 2D stage
    // of batch b.  Pinned caller memory (utr_host_alloc) is copied by DMA
    // directly; pageable memory is first copied by several host threads into
    // one of three pinned slots, while the DMA and kernels of the previous
    // batches run.
    const bool pinned = host_pinned(images, sizeof(float) * n2 * count);
    UTR_HIP_TRY(c->img_stage.alloc(2 * slot));
    if (!pinned) UTR_HIP_TRY(c->hstage.ensure(slot, c->stream));
    hipEvent_t e0 = c->ev[6];
    UTR_HIP_TRY(hipEventRecord(e0, c->stream));
    for (int r = 0, b = 0; r < count; r += kStageRows, b++) {
        const int k = std::min(kStageRows, count - r), s = b & 1, hs = b % 3;
        const size_t bytes = sizeof(float) * n2 * k;
        const float* src = images + (size_t)r * n2;
        if (!pinned) {
            UTR_HIP_TRY(hipEventSynchronize(c->hstage.ev[hs]));  // the slot's last DMA is done
            HostCopier::get().copy(c->hstage.at(hs), src, bytes);
            src = c->hstage.at(hs);
        }
        UTR_HIP_TRY(hipMemcpyAsync(c->img_stage.p + s * slot, src, bytes, hipMemcpyHostToDevice, c->stream));
        if (!pinned) UTR_HIP_TRY(hipEventRecord(c->hstage.ev[hs], c->stream));
        rc = enqueue_rows(c, first_row + r, k, c->img_stage.p + s * slot);
        if (rc) {
            (void)hipStreamSynchronize(c->stream);
            return rc;
        }
    }
    return finish_rows(c, e0, first_row, count);
}

int utr_host_alloc(size_t bytes, void** out) {
    if (!out) return set_err(UTR_E_INVALID, "null output pointer");
    *out = nullptr;
    void* p = nullptr;
    UTR_HIP_TRY(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault));
    {
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        g_pinned[(uintptr_t)p] = bytes ? bytes : 1;
    }
    *out = p;
    return UTR_OK;
}

int utr_host_free(void* p) {
    if (!p) return UTR_OK;
    {
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        auto it = g_pinned.find((uintptr_t)p);
        if (it == g_pinned.end()) return set_err(UTR_E_INVALID, "utr_host_free: not memory from utr_host_alloc");
        g_pinned.erase(it);
    }
    UTR_HIP_TRY(hipHostFree(p));
    return UTR_OK;
}

int utr_finalize_device(utr_ctx* c, float* d_vol) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (!d_vol) return set_err(UTR_E_INVALID, "null output pointer");
    return do_finalize(c, d_vol, nullptr);
}

int utr_finalize(utr_ctx* c, float* vol) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (!vol) return set_err(UTR_E_INVALID, "null output pointer");
    UTR_HIP_TRY(c->vol32.alloc(vox(c)));
    int rc = do_finalize(c, c->vol32.p, nullptr);
    if (rc) return rc;
    return copy_to_host(c, vol, c->vol32.p, vox(c));
}

int utr_finalize_f64(utr_ctx* c, double* vol) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (!vol) return set_err(UTR_E_INVALID, "null output pointer");
    if (c->precision != UTR_FP64) return set_err(UTR_E_INVALID, "utr_finalize_f64 needs a UTR_FP64 context");
    UTR_HIP_TRY(c->vol64.alloc(vox(c)));
    int rc = do_finalize(c, nullptr, c->vol64.p);
    if (rc) return rc;
    UTR_HIP_TRY(hipMemcpy(vol, c->vol64.p, sizeof(double) * vox(c), hipMemcpyDeviceToHost));
    return UTR_OK;
}

}  // extern "C"

namespace {

// Image correction table of rows from Z'' angles and shifts (either may be
// NULL: zeros), uploaded; for the DHT sCTF context, whose Z'' angles come
// after its creation (the UTR context takes them in utr_create).
int set_image_correction(utr_ctx* c, const double* angZ2, const double* shift_x, const double* shift_y) {
    if (c->n_added) return set_err(UTR_E_STATE, "image correction set after images were added");
    auto& P = c->plan;
    for (int i = 0; i < P.nrows; i++) {
        const double z2 = angZ2 ? angZ2[i] : 0.0, sx = shift_x ? shift_x[i] : 0.0,
                     sy = shift_y ? shift_y[i] : 0.0;
        if (!std::isfinite(z2) || !std::isfinite(sx) || !std::isfinite(sy))
            return set_err(UTR_E_INVALID, "image correction values must be finite");
        const double dx = std::floor(-sx / P.xst + 0.5), dy = std::floor(-sy / P.xst + 0.5);
        if (std::fabs(dx) > 1e6 || std::fabs(dy) > 1e6) return set_err(UTR_E_INVALID, "image shift too large");
        P.angZ2[i] = z2;
        const double a = -z2 * (3.14159265358979323846 / 180.0);
        utr::PrepRow& q = P.prep[i];
        q.cs = std::cos(a);
        q.sn = std::sin(a);
        q.dx = (int)dx;
        q.dy = (int)dy;
        q.on = (z2 != 0 || q.dx != 0 || q.dy != 0) ? 1 : 0;
    }
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    UTR_HIP_TRY(hipMemcpy(c->prep.p, P.prep.data(), sizeof(utr::PrepRow) * P.nrows, hipMemcpyHostToDevice));
    return alloc_prep(c);
}

}  // namespace

extern "C" {

int utr_set_image_shifts(utr_ctx* c, const double* shift_x, const double* shift_y) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (!shift_x || !shift_y) return set_err(UTR_E_INVALID, "null shift array");
    if (c->n_added) return set_err(UTR_E_STATE, "utr_set_image_shifts after images were added");
    auto& P = c->plan;
    for (int i = 0; i < P.nrows; i++) {
        if (!std::isfinite(shift_x[i]) || !std::isfinite(shift_y[i]))
            return set_err(UTR_E_INVALID, "image shifts must be finite");
        // UTR.cpp:965-969: Move(floor(-shift_y / yst + 0.5), floor(-shift_x / xst + 0.5))
        const double dx = std::floor(-shift_x[i] / P.xst + 0.5), dy = std::floor(-shift_y[i] / P.xst + 0.5);
        if (std::fabs(dx) > 1e6 || std::fabs(dy) > 1e6) return set_err(UTR_E_INVALID, "image shift too large");
        utr::PrepRow& q = P.prep[i];
        q.dx = (int)dx;
        q.dy = (int)dy;
        q.on = (P.angZ2[i] != 0 || q.dx != 0 || q.dy != 0) ? 1 : 0;
    }
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    UTR_HIP_TRY(hipMemcpy(c->prep.p, P.prep.data(), sizeof(utr::PrepRow) * P.nrows, hipMemcpyHostToDevice));
    return alloc_prep(c);
}

int utr_set_astigmatism(utr_ctx* c, const double* dz, const double* angle_deg) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (!dz || !angle_deg) return set_err(UTR_E_INVALID, "null astigmatism array");
    if (c->n_added) return set_err(UTR_E_STATE, "utr_set_astigmatism after images were added");
    auto& P = c->plan;
    for (int i = 0; i < P.nrows; i++)
        if (!std::isfinite(dz[i]) || !std::isfinite(angle_deg[i]))
            return set_err(UTR_E_INVALID, "astigmatism values must be finite");
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < P.nrows; i++) {
        // the restated mode-3 filter (UTR.cpp:1013): chi += pi wl dz ((qx^2 - qy^2) cos 2phi0 + 2 qx qy sin 2phi0)
        const double a2 = 2.0 * angle_deg[i] * pi / 180.0;
        P.rows[i].astig_c = pi * P.g.wl * dz[i] * std::cos(a2);
        P.rows[i].astig_s = pi * P.g.wl * dz[i] * std::sin(a2);
    }
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    UTR_HIP_TRY(hipMemcpy(c->rows.p, P.rows.data(), sizeof(utr::RowParams) * P.nrows, hipMemcpyHostToDevice));
    return UTR_OK;
}

int utr_keep_sampling(utr_ctx* c, int on) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    c->keep_sampling = on ? 1 : 0;
    return UTR_OK;
}

int utr_get_sampling(utr_ctx* c, float* samp) {
    if (!c || !samp) return set_err(UTR_E_INVALID, "null argument");
    if (!c->finalized) return set_err(UTR_E_STATE, "utr_get_sampling before utr_finalize");
    const size_t N = vox(c);
    if (c->precision != UTR_FP64) {
        if (!c->keep_sampling || !c->S32.p)
            return set_err(UTR_E_STATE, "sampling matrix not kept (call utr_keep_sampling before utr_finalize)");
        UTR_HIP_TRY(hipMemcpy(samp, c->S32.p, sizeof(float) * N, hipMemcpyDeviceToHost));
    } else {
        std::vector<double> tmp(N);
        UTR_HIP_TRY(hipMemcpy(tmp.data(), c->S64.p, sizeof(double) * N, hipMemcpyDeviceToHost));
        for (size_t t = 0; t < N; t++) samp[t] = (float)tmp[t];
    }
    return UTR_OK;
}

int utr_get_spectrum(utr_ctx* c, int row, float* spec) {
    if (!c || !spec) return set_err(UTR_E_INVALID, "null argument");
    if (row < 0 || row >= c->plan.nrows) return set_err(UTR_E_INVALID, "row outside the table");
    if (c->precision == UTR_FP64) return set_err(UTR_E_INVALID, "utr_get_spectrum needs an FP32 context");
    if (!c->added[row]) return set_err(UTR_E_STATE, "row has no image yet");
    const size_t n2 = (size_t)c->plan.g.nx * c->plan.g.ny;
    UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    UTR_HIP_TRY(hipMemcpy(spec, c->spec32.p + (size_t)row * n2, sizeof(cplx<float>) * n2,
                          hipMemcpyDeviceToHost));
    return UTR_OK;
}

int utr_get_fourier_volume(utr_ctx* c, float* out) {
    if (!c || !out) return set_err(UTR_E_INVALID, "null argument");
    if (!c->finalized) return set_err(UTR_E_STATE, "utr_get_fourier_volume before utr_finalize");
    const size_t N = vox(c);
    if (c->precision != UTR_FP64) {
        UTR_HIP_TRY(hipMemcpy(out, c->V32.p, sizeof(cplx<float>) * N, hipMemcpyDeviceToHost));
    } else {
        std::vector<cplx<double>> tmp(N);
        UTR_HIP_TRY(hipMemcpy(tmp.data(), c->V64.p, sizeof(cplx<double>) * N, hipMemcpyDeviceToHost));
        for (size_t t = 0; t < N; t++) {
            out[2 * t] = (float)tmp[t].x;
            out[2 * t + 1] = (float)tmp[t].y;
        }
    }
    return UTR_OK;
}

int utr_reset(utr_ctx* c) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    std::fill(c->added.begin(), c->added.end(), 0);
    c->n_added = 0;
    c->add_ms = 0;
    c->finalized = false;
    if (c->precision == UTR_FP64) {
        UTR_HIP_TRY(hipMemsetAsync(c->V64.p, 0, sizeof(cplx<double>) * vox(c), c->stream));
        UTR_HIP_TRY(hipMemsetAsync(c->S64.p, 0, sizeof(double) * vox(c), c->stream));
        UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
        UTR_HIP_TRY(hipMemsetAsync(c->rowmax.p, 0, sizeof(unsigned) * c->plan.nrows, c->stream));
    }
    return UTR_OK;
}

int utr_dims(const utr_ctx* c, int* nx, int* ny, int* nz) {
    if (!c) return set_err(UTR_E_INVALID, "null context");
    if (nx) *nx = c->plan.g.nx;
    if (ny) *ny = c->plan.g.ny;
    if (nz) *nz = c->plan.g.nz;
    return UTR_OK;
}

int utr_get_timings(const utr_ctx* c, utr_timings* t) {
    if (!c || !t) return set_err(UTR_E_INVALID, "null argument");
    *t = c->tim;
    return UTR_OK;
}

int utr_destroy(utr_ctx* c) {
    if (!c) return UTR_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->rows.release();
    c->img_stage.release();
    c->hstage.release();
    c->prep.release();
    c->img_prep.release();
    c->prep_aver.release();
    c->spec32.release(); c->V32.release(); c->H32.release(); c->S32.release(); c->vol32.release();
    c->twx32.release(); c->twy32.release(); c->twz32.release(); c->rowmax.release();
    c->tile_order.release(); c->rowc.release();
    c->spec64.release(); c->V64.release(); c->H64.release(); c->S64.release(); c->vol64.release();
    c->twx64.release(); c->twy64.release(); c->twz64.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return UTR_OK;
}

}  // extern "C"

// ---------------------------------------------------------- batched form


#include <tuple>

struct utr_batch {
    std::vector<utr_job> jobs;
    std::vector<utr_ctx*> ctx;
    std::vector<DevBuf<float>> img, vol;  // device copies for host-side jobs
    std::vector<const float*> d_img;
    std::vector<float*> d_vol;
    std::vector<hipEvent_t> join;
    hipStream_t s0 = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// The launches of one whole reconstruction on the context's stream, without
// host synchronisation or memset nodes (capturable).
int enqueue_job(utr_ctx* c, const float* d_img, float* d_vol) {
    const auto& g = c->plan.g;
    const int nrows = c->plan.nrows;
    utr::k_fill_u32<<<(nrows + 255) / 256, 256, 0, c->stream>>>(c->rowmax.p, nrows);
    const size_t n2 = (size_t)g.nx * g.ny;
    int rc = with_prep(c, 0, nrows, d_img, [&](int f, int k, const float* im) {
        return launch_2d<float>(c, im, f, k, c->spec32.p + (size_t)f * n2, c->twx32.p, c->twy32.p);
    });
    if (rc) return rc;
    if (c->precision == UTR_FP32) {
        rc = launch_grid32(c, nullptr, nullptr);
        if (rc) return rc;
    } else {
        const long nblocks = (long)((g.nx + utr::GBX - 1) / utr::GBX) *
                             ((g.ny + utr::GBY - 1) / utr::GBY) * ((g.nz + utr::GBZ - 1) / utr::GBZ);
        utr::k_gather<<<(unsigned)nblocks, 256, 0, c->stream>>>(c->spec32.p, c->rows.p, nrows, g,
                                                               c->V32.p, nullptr);
    }
    UTR_HIP_TRY(hipGetLastError());
    return launch_3d<float, float>(c, c->V32.p, c->H32.p, d_vol, c->twx32.p, c->twy32.p, c->twz32.p);
}

// The batched form's tuning of one job shape: gridding form, planes per
// plane-gridding workgroup, resident workgroups per CU of the persistent
// row and column FFTs.
struct JobTune {
    int form = UTR_FP32, planes = 0, rows_wpc = 0, cols_wpc = 0;
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, JobTune> g_tune;  // (nx, ny, nz, rows) -> tuning

int time_job(utr_ctx* c, const float* d_img, float* d_vol, float* ms) {
    hipEvent_t a = c->ev[4], b = c->ev[5];
    float best = 1e30f;
    for (int rep = 0; rep < 2; rep++) {
        UTR_HIP_TRY(hipEventRecord(a, c->stream));
        int rc = enqueue_job(c, d_img, d_vol);
        if (rc) return rc;
        UTR_HIP_TRY(hipEventRecord(b, c->stream));
        UTR_HIP_TRY(hipEventSynchronize(b));
        float t = 0;
        UTR_HIP_TRY(hipEventElapsedTime(&t, a, b));
        best = std::min(best, t);
    }
    *ms = best;
    return UTR_OK;
}

}  // namespace

extern "C" {

int utr_batch_destroy(utr_batch* b) {
    if (!b) return UTR_OK;
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto* c : b->ctx) utr_destroy(c);
    for (auto& d : b->img) d.release();
    for (auto& d : b->vol) d.release();
    for (auto e : b->join)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->s0) (void)hipStreamDestroy(b->s0);
    delete b;
    return UTR_OK;
}

int utr_batch_create(const utr_job* jobs, int njobs, int autotune, utr_batch** out) {
    if (!out) return set_err(UTR_E_INVALID, "null output pointer");
    *out = nullptr;
    if (!jobs || njobs < 1) return set_err(UTR_E_INVALID, "no jobs");
    utr_batch* b = new utr_batch();
    auto fail = [&](int rc) {
        utr_batch_destroy(b);
        return rc;
    };
    b->jobs.assign(jobs, jobs + njobs);
    b->ctx.assign(njobs, nullptr);
    b->img.resize(njobs);
    b->vol.resize(njobs);
    b->d_img.assign(njobs, nullptr);
    b->d_vol.assign(njobs, nullptr);
    b->join.assign(njobs, nullptr);
    if (hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->fork) != hipSuccess || hipEventCreate(&b->t0) != hipSuccess ||
        hipEventCreate(&b->t1) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "stream/event creation failed"));
    for (int j = 0; j < njobs; j++) {
        const utr_job& J = jobs[j];
        if (!J.images || !J.vol) return fail(set_err(UTR_E_INVALID, "job " + std::to_string(j) + ": null buffer"));
        const int form = J.precision == UTR_FP32_GATHER ? UTR_FP32_GATHER : UTR_FP32;
        int rc = utr_create(&J.params, J.nrows, J.angZ_deg, J.angY_deg, nullptr, J.defocus, J.weights,
                            form, &b->ctx[j]);
        if (rc) return fail(set_err(rc, "job " + std::to_string(j) + ": " + g_err));
        utr_ctx* c = b->ctx[j];
        const auto& g = c->plan.g;
        const size_t n2 = (size_t)g.nx * g.ny, N = n2 * g.nz;
        if (J.images_on_device) {
            b->d_img[j] = J.images;
        } else {
            if (b->img[j].alloc(n2 * J.nrows) != hipSuccess) return fail(set_err(UTR_E_NOMEM, "image buffer"));
            b->d_img[j] = b->img[j].p;
            if (hipMemcpy(b->img[j].p, J.images, sizeof(float) * n2 * J.nrows, hipMemcpyHostToDevice) != hipSuccess)
                return fail(set_err(UTR_E_DEVICE, "image upload"));
        }
        if (J.vol_on_device) {
            b->d_vol[j] = J.vol;
        } else {
            if (b->vol[j].alloc(N) != hipSuccess) return fail(set_err(UTR_E_NOMEM, "volume buffer"));
            b->d_vol[j] = b->vol[j].p;
        }
        if (hipEventCreateWithFlags(&b->join[j], hipEventDisableTiming) != hipSuccess)
            return fail(set_err(UTR_E_DEVICE, "event creation failed"));
        if (autotune) {
            const auto key = std::make_tuple(g.nx, g.ny, g.nz, J.nrows);
            bool have = false;
            JobTune T;
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tune.find(key);
                if (it != g_tune.end()) {
                    T = it->second;
                    have = true;
                }
            }
            if (!have) {
                // coordinate descent from the defaults, each candidate timed
                // on the job itself (best of two runs): planes per gridding
                // workgroup, the two FFT launch shapes, then the gather form
                auto apply = [&](const JobTune& t) {
                    c->precision = t.form;
                    c->grid_planes = t.planes;
                    c->rows_wpc = t.rows_wpc;
                    c->cols_wpc = t.cols_wpc;
                };
                float best = 0;
                apply(T);
                if ((rc = time_job(c, b->d_img[j], b->d_vol[j], &best))) return fail(rc);
                auto trial = [&](JobTune t) -> int {
                    apply(t);
                    float ms = 0;
                    const int r2 = time_job(c, b->d_img[j], b->d_vol[j], &ms);
                    if (r2) return r2;
                    if (ms < best) {
                        best = ms;
                        T = t;
                    }
                    apply(T);
                    return UTR_OK;
                };
                if (c->plan.all_y_exact)
                    for (int pl : {1, 2, 4}) {
                        JobTune t = T;
                        t.planes = pl;
                        if ((rc = trial(t))) return fail(rc);
                    }
                for (int rw : {4, 8}) {
                    JobTune t = T;
                    t.rows_wpc = rw;
                    if ((rc = trial(t))) return fail(rc);
                }
                for (int cw : {2, 4}) {
                    JobTune t = T;
                    t.cols_wpc = cw;
                    if ((rc = trial(t))) return fail(rc);
                }
                {
                    JobTune t = T;
                    t.form = UTR_FP32_GATHER;
                    if ((rc = trial(t))) return fail(rc);
                }
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tune[key] = T;
            }
            c->precision = T.form;
            c->grid_planes = T.planes;
            c->rows_wpc = T.rows_wpc;
            c->cols_wpc = T.cols_wpc;
        }
    }
    // capture: fork from s0, one branch per job on its own stream, join on s0
    for (auto* c : b->ctx) UTR_HIP_TRY(hipStreamSynchronize(c->stream));
    UTR_HIP_TRY(hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal));
    int rc = UTR_OK;
    if (hipEventRecord(b->fork, b->s0) != hipSuccess) rc = set_err(UTR_E_DEVICE, "capture: fork");
    for (int j = 0; j < njobs && !rc; j++) {
        utr_ctx* c = b->ctx[j];
        if (hipStreamWaitEvent(c->stream, b->fork, 0) != hipSuccess) {
            rc = set_err(UTR_E_DEVICE, "capture: wait");
            break;
        }
        rc = enqueue_job(c, b->d_img[j], b->d_vol[j]);
        if (rc) break;
        if (hipEventRecord(b->join[j], c->stream) != hipSuccess ||
            hipStreamWaitEvent(b->s0, b->join[j], 0) != hipSuccess)
            rc = set_err(UTR_E_DEVICE, "capture: join");
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(b->s0, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(rc);
    }
    if (e != hipSuccess) return fail(set_err(UTR_E_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(e)));
    b->graph = graph;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
        return fail(set_err(UTR_E_DEVICE, "graph instantiation failed"));
    *out = b;
    return UTR_OK;
}

int utr_batch_run(utr_batch* b, double* graph_ms) {
    if (!b) return set_err(UTR_E_INVALID, "null batch");
    const int njobs = (int)b->jobs.size();
    for (int j = 0; j < njobs; j++) {
        const utr_job& J = b->jobs[j];
        if (J.images_on_device) continue;
        const auto& g = b->ctx[j]->plan.g;
        UTR_HIP_TRY(hipMemcpyAsync(b->img[j].p, J.images, sizeof(float) * (size_t)g.nx * g.ny * J.nrows,
                                   hipMemcpyHostToDevice, b->s0));
    }
    UTR_HIP_TRY(hipEventRecord(b->t0, b->s0));
    UTR_HIP_TRY(hipGraphLaunch(b->exec, b->s0));
    UTR_HIP_TRY(hipEventRecord(b->t1, b->s0));
    for (int j = 0; j < njobs; j++) {
        const utr_job& J = b->jobs[j];
        if (J.vol_on_device) continue;
        const auto& g = b->ctx[j]->plan.g;
        UTR_HIP_TRY(hipMemcpyAsync(J.vol, b->vol[j].p, sizeof(float) * (size_t)g.nx * g.ny * g.nz,
                                   hipMemcpyDeviceToHost, b->s0));
    }
    UTR_HIP_TRY(hipStreamSynchronize(b->s0));
    if (graph_ms) {
        float ms = 0;
        UTR_HIP_TRY(hipEventElapsedTime(&ms, b->t0, b->t1));
        *graph_ms = ms;
    }
    return UTR_OK;
}

int utr_batch_launch(utr_batch* b, int repeats, double* ms) {
    if (!b) return set_err(UTR_E_INVALID, "null batch");
    if (repeats < 1) return set_err(UTR_E_INVALID, "repeats must be positive");
    for (const utr_job& J : b->jobs)
        if (!J.images_on_device || !J.vol_on_device)
            return set_err(UTR_E_INVALID, "utr_batch_launch: every job must have device images and volumes");
    UTR_HIP_TRY(hipEventRecord(b->t0, b->s0));
    for (int r = 0; r < repeats; r++) UTR_HIP_TRY(hipGraphLaunch(b->exec, b->s0));
    UTR_HIP_TRY(hipEventRecord(b->t1, b->s0));
    UTR_HIP_TRY(hipEventSynchronize(b->t1));
    if (ms) {
        float t = 0;
        UTR_HIP_TRY(hipEventElapsedTime(&t, b->t0, b->t1));
        *ms = t;
    }
    return UTR_OK;
}

int utr_batch_job_form(const utr_batch* b, int j) {
    if (!b || j < 0 || j >= (int)b->ctx.size()) return set_err(UTR_E_INVALID, "no such job");
    return b->ctx[j]->precision;
}

int utr_batch_job_tuning(const utr_batch* b, int j, int* form, int* planes, int* rows_wpc, int* cols_wpc) {
    if (!b || j < 0 || j >= (int)b->ctx.size()) return set_err(UTR_E_INVALID, "no such job");
    const utr_ctx* c = b->ctx[j];
    if (form) *form = c->precision;
    if (planes) *planes = c->grid_planes;
    if (rows_wpc) *rows_wpc = c->rows_wpc;
    if (cols_wpc) *cols_wpc = c->cols_wpc;
    return UTR_OK;
}

}  // extern "C"

// DHT variants (methods 1-5) share the helpers above
#include "dht.hip"
#include "dht_series.hip"
// post-processing of the reconstructed volume
#include "post.hip"

#include "msct.hip"

#include "evaluate.hip"

#include "corrector.hip"
