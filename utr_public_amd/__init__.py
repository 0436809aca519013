"""utr_public_amd -- MI355X-native UTR (Unified Tomographic Reconstruction).

Python bindings (ctypes) over the C ABI of ``_lib/libutr_hip.so``
(``include/utr_hip.h``).  The library runs hand-written gfx950 HIP kernels;
there is no CPU fallback: if the library is missing, ``load()`` raises.

Reference: timg021/UTR_public, MultisliceCppTest/PhaseRetrieval/UTR.cpp.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libutr_hip.so")
SYNTH_PATH = os.path.join(_PKG, "_lib", "libutr_synth.so")

UTR_FP32 = 0          # brick scatter, exact fixed-point sums (default)
UTR_FP64 = 1          # plain FP64 form
UTR_FP32_GATHER = 2   # per-voxel gather, Kahan-compensated FP32 sums

PARAM_FIELDS = [
    ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("modality", ctypes.c_int),
    ("xst", ctypes.c_double), ("zlo", ctypes.c_double), ("zhi", ctypes.c_double),
    ("dzextra", ctypes.c_double), ("dxc", ctypes.c_double), ("dyc", ctypes.c_double),
    ("dzc", ctypes.c_double), ("wl", ctypes.c_double), ("aobj_mrad", ctypes.c_double),
    ("cs3_mm", ctypes.c_double), ("cs5_mm", ctypes.c_double), ("ctf_mode", ctypes.c_int),
    ("escc", ctypes.c_int), ("asigma", ctypes.c_double), ("nsr", ctypes.c_double),
    ("epsilon", ctypes.c_double), ("inv_laplace", ctypes.c_int),
    ("norm_f1", ctypes.c_double), ("norm_f2", ctypes.c_double),
]

# Symbols include/utr_hip.h declares (checked by the CPU test suite).
EXPORTS = [
    "utr_create", "utr_add_images", "utr_add_images_device", "utr_finalize",
    "utr_finalize_device", "utr_finalize_f64", "utr_keep_sampling", "utr_get_sampling",
    "utr_set_image_shifts", "utr_set_astigmatism",
    "utr_reset", "utr_dims", "utr_get_timings", "utr_destroy", "utr_last_error",
    "utr_abi_version", "utr_angle_coeffs", "utr_get_spectrum", "utr_get_fourier_volume",
    "utr_batch_create", "utr_batch_run", "utr_batch_launch", "utr_batch_job_form", "utr_batch_job_tuning",
    "utr_batch_destroy",
    "dht_create", "dht_add_images", "dht_add_images_device", "dht_finalize", "dht_finalize_device",
    "dht_reset", "dht_dims", "dht_get_timings", "dht_destroy", "dht_set_image_correction",
    "dht_create_series", "dht_get_iterations", "dht_get_waves", "utr_host_alloc", "utr_host_free",
    "utr_highpass3d", "utr_highpass3d_device", "utr_highpass3d_device_noalloc", "utr_rotate3d", "utr_rotate3d_device",
    "utr_symmetrize3d", "utr_symmetrize3d_device", "utr_invlaplace3d", "utr_invlaplace3d_device",
    "msct_create", "msct_images", "msct_images_device", "msct_slices", "msct_batch", "msct_destroy",
    "utr_ground_truth_device", "utr_reproject_device",
    "utr_threshold", "utr_threshold_device", "utr_reslice", "utr_reslice_device", "utr_thick_slices",
    "utr_thick_slices_device", "utr_convert16", "utr_convert16_device",
    "utr_corr_create", "utr_corr_nparams", "utr_corr_tile", "utr_corr_get_params", "utr_corr_set_params",
    "utr_corr_forward", "utr_corr_gradient", "utr_corr_train_step", "utr_corr_apply", "utr_corr_destroy",
]

DHT_FIELDS = [
    ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("xst", ctypes.c_double),
    ("zlo", ctypes.c_double), ("zhi", ctypes.c_double), ("dzextra", ctypes.c_double),
    ("dxc", ctypes.c_double), ("dyc", ctypes.c_double), ("dzc", ctypes.c_double),
    ("wl", ctypes.c_double), ("aobj_mrad", ctypes.c_double), ("cs3_mm", ctypes.c_double),
    ("cs5_mm", ctypes.c_double), ("asigma", ctypes.c_double), ("method", ctypes.c_int),
    ("escc", ctypes.c_int), ("critical_snr", ctypes.c_double), ("epsilon", ctypes.c_double),
    ("inv_laplace", ctypes.c_int), ("alpha", ctypes.c_double),
    ("norm_f1", ctypes.c_double), ("norm_f2", ctypes.c_double), ("itermax", ctypes.c_int),
]


MSCT_FIELDS = [
    ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("mode", ctypes.c_int), ("output", ctypes.c_int),
    ("ctblength", ctypes.c_double), ("deltaz", ctypes.c_double), ("kev", ctypes.c_double),
    ("aobj_mrad", ctypes.c_double), ("cs3_mm", ctypes.c_double), ("cs5_mm", ctypes.c_double),
    ("astig_dz", ctypes.c_double), ("astig_angle_deg", ctypes.c_double),
    ("dxc", ctypes.c_double), ("dyc", ctypes.c_double), ("dzc", ctypes.c_double),
    ("nwobble", ctypes.c_int), ("reserved", ctypes.c_int), ("temperature", ctypes.c_double),
    ("seed", ctypes.c_ulonglong),
]


class MsctParams(ctypes.Structure):
    _fields_ = MSCT_FIELDS


class UtrParams(ctypes.Structure):
    _fields_ = PARAM_FIELDS


class DhtParams(ctypes.Structure):
    _fields_ = DHT_FIELDS


class UtrJob(ctypes.Structure):
    _fields_ = [("params", UtrParams), ("nrows", ctypes.c_int),
                ("angZ_deg", ctypes.c_void_p), ("angY_deg", ctypes.c_void_p),
                ("defocus", ctypes.c_void_p), ("weights", ctypes.c_void_p),
                ("images", ctypes.c_void_p), ("vol", ctypes.c_void_p),
                ("images_on_device", ctypes.c_int), ("vol_on_device", ctypes.c_int),
                ("precision", ctypes.c_int)]


class UtrTimings(ctypes.Structure):
    _fields_ = [("fft2d_ctf_ms", ctypes.c_double), ("backprop_ms", ctypes.c_double),
                ("fft3d_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class UTRError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("UTR error %d: %s" % (code, msg))
        self.code = code


def default_params(**kw):
    """Parameter set with the reference's neutral defaults (UTR.txt)."""
    d = dict(nx=128, ny=128, modality=0, xst=0.4, zlo=-25.6, zhi=25.6, dzextra=0.0,
             dxc=0.0, dyc=0.0, dzc=0.0, wl=0.02507934, aobj_mrad=0.0, cs3_mm=0.0,
             cs5_mm=0.0, ctf_mode=0, escc=1, asigma=0.0, nsr=0.0, epsilon=1e-4,
             inv_laplace=0, norm_f1=1.0, norm_f2=0.0)
    d.update(kw)
    return d


def default_dht_params(**kw):
    """DHT parameter set: method 3 (Min05LogAmp + Fresnel back-propagation)
    on a 64^2 x 64 grid at 0.4 A, 200 keV, 100 A extra defocus."""
    d = dict(nx=64, ny=64, xst=0.4, zlo=-12.8, zhi=12.8, dzextra=100.0, dxc=0.0, dyc=0.0,
             dzc=0.0, wl=0.02507934, aobj_mrad=0.0, cs3_mm=0.0, cs5_mm=0.0, asigma=0.0,
             method=3, escc=1, critical_snr=0.5, epsilon=1.0, inv_laplace=0, alpha=0.0,
             norm_f1=1.0, norm_f2=0.0, itermax=200)
    d.update(kw)
    return d


def make_params(d):
    p = UtrParams()
    for name, _ in PARAM_FIELDS:
        setattr(p, name, d[name])
    return p


_lib = None


def load():
    """Loads libutr_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libutr_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Loading torch first makes this library bind to that already-loaded
    # runtime, so torch tensors and this library share one HIP runtime in the
    # process; loaded the other way round, torch would start a second runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    L.utr_create.argtypes = [ctypes.POINTER(UtrParams), ctypes.c_int, dp, dp, dp, dp, vp,
                             ctypes.c_int, ctypes.POINTER(vp)]
    L.utr_add_images.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.utr_add_images_device.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.utr_finalize.argtypes = [vp, vp]
    L.utr_finalize_device.argtypes = [vp, vp]
    L.utr_finalize_f64.argtypes = [vp, vp]
    L.utr_keep_sampling.argtypes = [vp, ctypes.c_int]
    L.utr_set_image_shifts.argtypes = [vp, dp, dp]
    L.utr_set_astigmatism.argtypes = [vp, dp, dp]
    L.utr_get_sampling.argtypes = [vp, vp]
    L.utr_reset.argtypes = [vp]
    L.utr_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(ctypes.c_int)]
    L.utr_get_timings.argtypes = [vp, ctypes.POINTER(UtrTimings)]
    L.utr_destroy.argtypes = [vp]
    L.utr_last_error.restype = ctypes.c_char_p
    L.utr_angle_coeffs.argtypes = [ctypes.POINTER(UtrParams), ctypes.c_double, ctypes.c_double, dp]
    L.utr_get_spectrum.argtypes = [vp, ctypes.c_int, vp]
    L.utr_get_fourier_volume.argtypes = [vp, vp]
    L.utr_batch_create.argtypes = [ctypes.POINTER(UtrJob), ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.utr_batch_run.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.utr_batch_launch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.utr_batch_job_form.argtypes = [vp, ctypes.c_int]
    ip = ctypes.POINTER(ctypes.c_int)
    L.utr_batch_job_tuning.argtypes = [vp, ctypes.c_int, ip, ip, ip, ip]
    L.utr_batch_destroy.argtypes = [vp]
    L.dht_create.argtypes = [ctypes.POINTER(DhtParams), ctypes.c_int, dp, dp, dp, vp,
                             ctypes.POINTER(vp)]
    for f in ("dht_add_images", "dht_add_images_device"):
        getattr(L, f).argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.dht_finalize.argtypes = [vp, vp]
    L.dht_finalize_device.argtypes = [vp, vp]
    L.dht_reset.argtypes = [vp]
    L.dht_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(ctypes.c_int)]
    L.dht_get_timings.argtypes = [vp, ctypes.POINTER(UtrTimings)]
    L.dht_destroy.argtypes = [vp]
    L.dht_set_image_correction.argtypes = [vp, dp, dp, dp]
    L.utr_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
    L.utr_host_free.argtypes = [vp]
    L.dht_create_series.argtypes = [ctypes.POINTER(DhtParams), ctypes.c_int, dp, dp,
                                    ctypes.POINTER(ctypes.c_int), dp, vp, ctypes.POINTER(vp)]
    L.dht_get_iterations.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.dht_get_waves.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    d = ctypes.c_double
    i = ctypes.c_int
    for f in ("utr_highpass3d", "utr_highpass3d_device", "utr_highpass3d_device_noalloc"):
        getattr(L, f).argtypes = [vp, i, i, i, d, d, d, d]
    for f in ("utr_invlaplace3d", "utr_invlaplace3d_device"):
        getattr(L, f).argtypes = [vp, i, i, i, d, d, d, d, d]
    L.utr_rotate3d.argtypes = [vp, i, i, i, d, d, d]
    L.utr_rotate3d_device.argtypes = [vp, vp, i, i, i, d, d, d]
    for f in ("utr_symmetrize3d", "utr_symmetrize3d_device"):
        getattr(L, f).argtypes = [vp, i, i, i, i, dp, dp, dp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.msct_create.argtypes = [ctypes.POINTER(MsctParams), dp, i, ctypes.POINTER(ctypes.c_int), fp, fp, fp,
                              fp, fp, ctypes.POINTER(vp)]
    L.msct_images.argtypes = [vp, i, dp, dp, i, dp, vp]
    L.msct_images_device.argtypes = [vp, i, dp, dp, i, dp, vp]
    L.msct_slices.argtypes = [vp]
    L.msct_batch.argtypes = [vp]
    L.msct_destroy.argtypes = [vp]
    fp_ = ctypes.POINTER(ctypes.c_float)
    ip_ = ctypes.POINTER(ctypes.c_int)
    ll = ctypes.c_longlong
    for f in ("utr_threshold", "utr_threshold_device"):
        getattr(L, f).argtypes = [vp, ll, ctypes.c_float, ctypes.c_float]
    for f in ("utr_reslice", "utr_reslice_device"):
        getattr(L, f).argtypes = [vp, vp, i, i, i, i, i, i, i]
    for f in ("utr_thick_slices", "utr_thick_slices_device"):
        getattr(L, f).argtypes = [vp, vp, i, i, i, i, i]
    for f in ("utr_convert16", "utr_convert16_device"):
        getattr(L, f).argtypes = [vp, vp, ll, ctypes.c_float, ctypes.c_float, ll]
    L.utr_ground_truth_device.argtypes = [i, i, i, d, d, d, dp, i, ip_, fp_, fp_, fp_, vp]
    L.utr_reproject_device.argtypes = [ctypes.POINTER(UtrParams), vp, i, dp, dp, dp, vp, vp, dp]
    L.utr_corr_create.argtypes = [i, i, i, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.utr_corr_nparams.argtypes = [vp]
    L.utr_corr_tile.argtypes = [vp, ip_, ip_, ip_]
    L.utr_corr_get_params.argtypes = [vp, fp_, fp_]
    L.utr_corr_set_params.argtypes = [vp, fp_, ctypes.c_float]
    L.utr_corr_forward.argtypes = [vp, vp, i, vp]
    L.utr_corr_gradient.argtypes = [vp, vp, vp, i, fp_, fp_]
    L.utr_corr_train_step.argtypes = [vp, vp, vp, i, ctypes.c_float, fp_]
    L.utr_corr_apply.argtypes = [vp, vp, vp, i, i, i]
    L.utr_corr_destroy.argtypes = [vp]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise UTRError(rc, load().utr_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def angle_coeffs(params, angZ_deg, angY_deg):
    m = np.zeros(9)
    p = make_params(params)
    _check(load().utr_angle_coeffs(ctypes.byref(p), angZ_deg, angY_deg, _dp(m)))
    return m.reshape(3, 3)


class Reconstructor:
    """One UTR reconstruction context on the current HIP device.

    rows: dict with arrays angZ, angY (degrees), defocus (UL), optional angZ2
    and weights -- the columns of the orientation file read at UTR.cpp:122 --
    and optional shift_x, shift_y (UL, the image shifts of a .relionnew file)
    and astig_dz, astig_angle (its astigmatism, used by ctf_mode 3).
    """

    def __init__(self, params, rows, precision=UTR_FP32):
        L = load()
        self._L = L
        self.params = dict(params)
        self.angZ = np.ascontiguousarray(rows["angZ"], dtype=np.float64)
        self.angY = np.ascontiguousarray(rows["angY"], dtype=np.float64)
        self.defocus = np.ascontiguousarray(rows["defocus"], dtype=np.float64)
        self.angZ2 = None
        if rows.get("angZ2") is not None:
            self.angZ2 = np.ascontiguousarray(rows["angZ2"], dtype=np.float64)
        self.weights = None
        if rows.get("weights") is not None:
            self.weights = np.ascontiguousarray(rows["weights"], dtype=np.float32)
        self.nrows = len(self.angZ)
        self.precision = precision
        p = make_params(self.params)
        ctx = ctypes.c_void_p()
        _check(L.utr_create(ctypes.byref(p), self.nrows, _dp(self.angZ), _dp(self.angY),
                            _dp(self.angZ2) if self.angZ2 is not None else None,
                            _dp(self.defocus),
                            self.weights.ctypes.data if self.weights is not None else None,
                            precision, ctypes.byref(ctx)))
        self._ctx = ctx
        if rows.get("shift_x") is not None or rows.get("shift_y") is not None:
            z = np.zeros(self.nrows)
            sx = np.ascontiguousarray(rows.get("shift_x", z), dtype=np.float64)
            sy = np.ascontiguousarray(rows.get("shift_y", z), dtype=np.float64)
            assert sx.shape == (self.nrows,) and sy.shape == (self.nrows,)
            _check(L.utr_set_image_shifts(ctx, _dp(sx), _dp(sy)))
        if rows.get("astig_dz") is not None:
            adz = np.ascontiguousarray(rows["astig_dz"], dtype=np.float64)
            aang = np.ascontiguousarray(rows.get("astig_angle", np.zeros(self.nrows)), dtype=np.float64)
            assert adz.shape == (self.nrows,) and aang.shape == (self.nrows,)
            _check(L.utr_set_astigmatism(ctx, _dp(adz), _dp(aang)))
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(L.utr_dims(ctx, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        self.shape = (nz.value, ny.value, nx.value)

    def add_images(self, images, first_row=0):
        images = np.ascontiguousarray(images, dtype=np.float32)
        n = images.shape[0]
        assert images.shape[1:] == self.shape[1:], (images.shape, self.shape)
        _check(self._L.utr_add_images(self._ctx, first_row, n, images.ctypes.data))

    def add_images_device(self, ptr, count, first_row=0):
        _check(self._L.utr_add_images_device(self._ctx, first_row, count, ctypes.c_void_p(ptr)))

    def keep_sampling(self, on=True):
        _check(self._L.utr_keep_sampling(self._ctx, 1 if on else 0))

    def finalize(self, out=None):
        if out is None:
            out = np.empty(self.shape, dtype=np.float32)
        assert out.dtype == np.float32 and out.shape == self.shape and out.flags.c_contiguous
        _check(self._L.utr_finalize(self._ctx, out.ctypes.data))
        return out

    def finalize_f64(self):
        out = np.empty(self.shape, dtype=np.float64)
        _check(self._L.utr_finalize_f64(self._ctx, out.ctypes.data))
        return out

    def finalize_device(self, ptr):
        _check(self._L.utr_finalize_device(self._ctx, ctypes.c_void_p(ptr)))

    def sampling(self):
        out = np.empty(self.shape, dtype=np.float32)
        _check(self._L.utr_get_sampling(self._ctx, out.ctypes.data))
        return out

    def spectrum(self, row):
        out = np.empty(self.shape[1:], dtype=np.complex64)
        _check(self._L.utr_get_spectrum(self._ctx, row, out.ctypes.data))
        return out

    def fourier_volume(self):
        out = np.empty(self.shape, dtype=np.complex64)
        _check(self._L.utr_get_fourier_volume(self._ctx, out.ctypes.data))
        return out

    def reset(self):
        _check(self._L.utr_reset(self._ctx))

    def timings(self):
        t = UtrTimings()
        _check(self._L.utr_get_timings(self._ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.utr_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def reconstruct(params, rows, images, precision=UTR_FP32, f64_out=False):
    """Convenience: one full reconstruction with host input and output."""
    with Reconstructor(params, rows, precision) as r:
        r.add_images(images)
        if f64_out:
            return r.finalize_f64()
        return r.finalize()


class Batch:
    """Independent reconstructions of uneven size run as one HIP graph
    (utr_batch_* of include/utr_hip.h).  jobs: list of (params, rows, images)
    with host images (numpy; volumes returned by run() as numpy arrays) or
    device images (torch CUDA float32 tensors; volumes allocated as torch
    CUDA tensors, and launch() replays the graph with no host copies)."""

    def __init__(self, jobs, autotune=False, precision=UTR_FP32):
        L = load()
        self._L = L
        self._keep = []
        self.shapes = []
        arr = (UtrJob * len(jobs))()
        self.outputs = []
        for k, (params, rows, images) in enumerate(jobs):
            J = arr[k]
            J.params = make_params(params)
            cols = [np.ascontiguousarray(rows[c], dtype=np.float64) for c in ("angZ", "angY", "defocus")]
            w = rows.get("weights")
            w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
            nz = int((max(params["zlo"], params["zhi"]) - min(params["zlo"], params["zhi"])) /
                     params["xst"] + 0.5)
            on_dev = hasattr(images, "is_cuda") and images.is_cuda
            if on_dev:
                import torch
                if images.dtype != torch.float32 or not images.is_contiguous():
                    raise ValueError("device images must be a contiguous float32 CUDA tensor")
                if images.numel() != len(cols[0]) * params["ny"] * params["nx"]:
                    raise ValueError("device images do not match the rows and image size")
                out = torch.empty((nz, params["ny"], params["nx"]), dtype=torch.float32, device=images.device)
                J.images = images.data_ptr()
                J.vol = out.data_ptr()
            else:
                images = np.ascontiguousarray(images, dtype=np.float32)
                out = np.empty((nz, params["ny"], params["nx"]), dtype=np.float32)
                J.images = images.ctypes.data
                J.vol = out.ctypes.data
            self._keep += cols + [w, images, out]
            J.nrows = len(cols[0])
            J.angZ_deg, J.angY_deg, J.defocus = (c.ctypes.data for c in cols)
            J.weights = w.ctypes.data if w is not None else None
            J.images_on_device = 1 if on_dev else 0
            J.vol_on_device = 1 if on_dev else 0
            J.precision = precision
            self.outputs.append(out)
        self._jobs = arr
        b = ctypes.c_void_p()
        _check(L.utr_batch_create(arr, len(jobs), 1 if autotune else 0, ctypes.byref(b)))
        self._b = b

    def run(self):
        ms = ctypes.c_double()
        _check(self._L.utr_batch_run(self._b, ctypes.byref(ms)))
        self.graph_ms = ms.value
        return self.outputs

    def launch(self, repeats=1):
        """Replays the graph `repeats` times back to back (device jobs only);
        returns the device milliseconds of all repeats."""
        ms = ctypes.c_double()
        _check(self._L.utr_batch_launch(self._b, int(repeats), ctypes.byref(ms)))
        return ms.value

    def form(self, j):
        return self._L.utr_batch_job_form(self._b, j)

    def tuning(self, j):
        """Launch configuration captured for job j: form, planes per gridding
        workgroup, row and column FFT workgroups per CU (0: default)."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(self._L.utr_batch_job_tuning(self._b, j, *[ctypes.byref(x) for x in v]))
        return dict(zip(("form", "planes", "rows_wpc", "cols_wpc"), (x.value for x in v)))

    def close(self):
        if getattr(self, "_b", None):
            self._L.utr_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DhtReconstructor:
    """One reconstruction of the DHT program (MultisliceCppTest/
    PhaseRetrieval/DHT.cpp) on the current HIP device.  rows: angZ, angY
    (degrees) and optional weights per orientation; defocus (Angstrom),
    angZ2 (Z'' in degrees) and shift_x, shift_y (RELION image shifts,
    Angstrom) per input image; counts (optional): images per orientation,
    the defocus series of methods 1 and 2 (dht_create_series)."""

    def __init__(self, params, rows):
        L = load()
        self._L = L
        self.params = dict(params)
        self.params.setdefault("itermax", 0)
        self.angZ = np.ascontiguousarray(rows["angZ"], dtype=np.float64)
        self.angY = np.ascontiguousarray(rows["angY"], dtype=np.float64)
        self.defocus = np.ascontiguousarray(rows["defocus"], dtype=np.float64)
        self.weights = None
        if rows.get("weights") is not None:
            self.weights = np.ascontiguousarray(rows["weights"], dtype=np.float32)
        self.counts = None
        if rows.get("counts") is not None:
            self.counts = np.ascontiguousarray(rows["counts"], dtype=np.int32)
            assert self.counts.shape == self.angZ.shape and self.counts.sum() == len(self.defocus)
        p = DhtParams()
        for name, _ in DHT_FIELDS:
            setattr(p, name, self.params[name])
        ctx = ctypes.c_void_p()
        w = self.weights.ctypes.data if self.weights is not None else None
        if self.counts is None:
            _check(L.dht_create(ctypes.byref(p), len(self.angZ), _dp(self.angZ), _dp(self.angY),
                                _dp(self.defocus), w, ctypes.byref(ctx)))
        else:
            _check(L.dht_create_series(ctypes.byref(p), len(self.angZ), _dp(self.angZ), _dp(self.angY),
                                       self.counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       _dp(self.defocus), w, ctypes.byref(ctx)))
        self._ctx = ctx
        if any(rows.get(k) is not None for k in ("angZ2", "shift_x", "shift_y")):
            z = np.zeros(len(self.defocus))
            cols = [np.ascontiguousarray(rows.get(k) if rows.get(k) is not None else z, dtype=np.float64)
                    for k in ("angZ2", "shift_x", "shift_y")]
            assert all(c.shape == z.shape for c in cols)
            _check(L.dht_set_image_correction(ctx, *(_dp(c) for c in cols)))
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(L.dht_dims(ctx, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        self.shape = (nz.value, ny.value, nx.value)

    def add_images(self, images, first_row=0):
        images = np.ascontiguousarray(images, dtype=np.float32)
        assert images.shape[1:] == self.shape[1:], (images.shape, self.shape)
        _check(self._L.dht_add_images(self._ctx, first_row, images.shape[0], images.ctypes.data))

    def add_images_device(self, ptr, count, first_row=0):
        _check(self._L.dht_add_images_device(self._ctx, first_row, count, ctypes.c_void_p(ptr)))

    def finalize(self):
        out = np.empty(self.shape, dtype=np.float32)
        _check(self._L.dht_finalize(self._ctx, out.ctypes.data))
        return out

    def finalize_device(self, ptr):
        _check(self._L.dht_finalize_device(self._ctx, ctypes.c_void_p(ptr)))

    def reset(self):
        _check(self._L.dht_reset(self._ctx))

    def timings(self):
        t = UtrTimings()
        _check(self._L.dht_get_timings(self._ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def iterations(self):
        """Methods 1, 2: IWFR iterations per orientation of the last finalize."""
        it = np.zeros(len(self.angZ), np.int32)
        _check(self._L.dht_get_iterations(self._ctx, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return it

    def waves(self, first=0, count=None):
        """Methods 1, 2, after finalize: the retrieved complex amplitudes
        (count, ny, nx) at each orientation's zout."""
        count = len(self.angZ) - first if count is None else count
        out = np.empty((count,) + self.shape[1:] + (2,), dtype=np.float32)
        _check(self._L.dht_get_waves(self._ctx, first, count, out.ctypes.data))
        return out[..., 0] + 1j * out[..., 1]

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.dht_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def highpass3d(vol, xst, yst, zst, width):
    """High-pass filter of a (nz, ny, nx) float32 volume on the GPU:
    vol - Gaussian(vol) with sigma = width / 2.355 (UTR.cpp:1357-1378).
    Returns a new array."""
    out = np.ascontiguousarray(vol, dtype=np.float32).copy()
    nz, ny, nx = out.shape
    _check(load().utr_highpass3d(out.ctypes.data, nx, ny, nz, xst, yst, zst, width))
    return out


def threshold(vol, background, threshold):
    """ThresholdLow of parameter 23 on the GPU (UTR.cpp:1383): a new array
    with vol - background, values below threshold set to it."""
    out = np.ascontiguousarray(vol, dtype=np.float32).copy()
    _check(load().utr_threshold(out.ctypes.data, out.size, float(background), float(threshold)))
    return out


_RESLICE_AXES = [(0, 1, 2), (0, 2, 1), (1, 2, 0), (1, 0, 2), (2, 0, 1), (2, 1, 0)]


def reslice(vol, mode, flipx=False, flipy=False, flipz=False):
    """Reslice with flips on the GPU (parameters 29-30, UTR.cpp:1449-1453) of
    a (nz, ny, nx) volume; mode 0 XY .. 5 ZY."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, nx = vol.shape
    n = (nx, ny, nz)
    a = _RESLICE_AXES[mode]
    out = np.empty((n[a[2]], n[a[1]], n[a[0]]), np.float32)
    _check(load().utr_reslice(vol.ctypes.data, out.ctypes.data, nx, ny, nz, mode, int(flipx), int(flipy),
                              int(flipz)))
    return out


def thick_slices(vol, mode, step):
    """Thick MIP (1) or average (2) slices over step slices on the GPU
    (parameter 31, UTR.cpp:1456-1490)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, nx = vol.shape
    out = np.empty((max(nz // max(step, 1), 0), ny, nx), np.float32)
    _check(load().utr_thick_slices(vol.ctypes.data, out.ctypes.data, nx, ny, nz, mode, step))
    return out


def convert16(vol, minin, maxin, maxout):
    """16-bit output on the GPU (parameter 27, UTR.cpp:1494-1498)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    out = np.empty(vol.shape, np.uint16)
    _check(load().utr_convert16(vol.ctypes.data, out.ctypes.data, vol.size, float(minin), float(maxin),
                                int(maxout)))
    return out


def invlaplace3d(vol, xst, yst, zst, alpha, norm):
    """Regularised inverse (-Laplacian) of a (nz, ny, nx) float32 volume on
    the GPU (UTR.cpp:1363-1367).  Returns a new array."""
    out = np.ascontiguousarray(vol, dtype=np.float32).copy()
    nz, ny, nx = out.shape
    _check(load().utr_invlaplace3d(out.ctypes.data, nx, ny, nz, xst, yst, zst, alpha, norm))
    return out


def rotate3d(vol, angZ, angY, angZ2):
    """3D rotation (degrees) of a (nz, ny, nx) float32 volume on the GPU: the
    content turns by Rz(angZ2) Ry(angY) Rz(angZ) about the array centre
    (Rotate3 of UTR.cpp:1393, parameter 28).  Returns a new array."""
    out = np.ascontiguousarray(vol, dtype=np.float32).copy()
    nz, ny, nx = out.shape
    _check(load().utr_rotate3d(out.ctypes.data, nx, ny, nz, float(angZ), float(angY), float(angZ2)))
    return out


def symmetrize3d(vol, angZ, angY, angZ2):
    """Symmetry average of UTR.cpp:1341-1355 (parameter 20 = 2) on the GPU;
    angZ, angY, angZ2: the rows of the symmetry file (degrees).  Returns a
    new array."""
    out = np.ascontiguousarray(vol, dtype=np.float32).copy()
    nz, ny, nx = out.shape
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (angZ, angY, angZ2)]
    assert a[0].shape == a[1].shape == a[2].shape and a[0].ndim == 1
    _check(load().utr_symmetrize3d(out.ctypes.data, nx, ny, nz, len(a[0]), *(_dp(x) for x in a)))
    return out


class _PinnedBlock:
    def __init__(self, nbytes):
        self._L = load()
        ptr = ctypes.c_void_p()
        _check(self._L.utr_host_alloc(nbytes, ctypes.byref(ptr)))
        self.ptr = ptr.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self._L.utr_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


def host_empty(shape, dtype=np.float32):
    """A numpy array in pinned host memory (utr_host_alloc): the library's
    host-memory entry points copy it by DMA without staging.  Freed with the
    array."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    owner = _PinnedBlock(nbytes)
    buf = (ctypes.c_char * max(nbytes, 1)).from_address(owner.ptr)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


def dht_reconstruct(params, rows, images):
    """Convenience: one DHT reconstruction with host input and output."""
    with DhtReconstructor(params, rows) as r:
        r.add_images(images)
        return r.finalize()


from . import synth  # noqa: E402,F401


# --------------------------------------------------------------------------
# TEM multislice forward model (msct_* C ABI)

FE_PARAMS_PATH = os.path.join(_PKG, "data", "fe_params.txt")


def fe_params():
    """Kirkland (2010) App. C electron scattering-factor parameters,
    (103, 12) for Z = 1..103 (the table of TemsimLibrary/slicelib.cpp:958-2209)."""
    return np.loadtxt(FE_PARAMS_PATH, dtype=np.float64).reshape(103, 12)


def default_msct_params(**kw):
    """Multislice set-up: 128^2 pixels over a 51.2 A cube, 1 A slices, 200 keV,
    no aperture or aberrations (the autosliccmd parameter lines)."""
    d = dict(nx=128, ny=128, mode=0, output=0, ctblength=51.2, deltaz=1.0, kev=200.0, aobj_mrad=0.0,
             cs3_mm=0.0, cs5_mm=0.0, astig_dz=0.0, astig_angle_deg=0.0, dxc=0.0, dyc=0.0, dzc=0.0,
             nwobble=0, reserved=0, temperature=300.0, seed=1)
    d.update(kw)
    return d


class Multislice:
    """One atom set for the multislice forward model (autosliccmd()):
    atoms = dict(Z, x, y, z[, occ][, wobble]) in Angstrom inside
    [0, ctblength]^3; wobble is the rms thermal displacement at 300 K, used
    when params nwobble >= 1 (1: Debye-Waller factor, > 1: configurations)."""

    def __init__(self, params, atoms):
        self._L = load()
        self.params = dict(params)
        p = MsctParams()
        for name, _ in MSCT_FIELDS:
            setattr(p, name, self.params[name])
        Z = np.ascontiguousarray(atoms["Z"], dtype=np.int32)
        n = len(Z)
        xs = [np.ascontiguousarray(atoms[k], dtype=np.float32) for k in ("x", "y", "z")]
        occ = np.ascontiguousarray(atoms.get("occ", np.ones(n)), dtype=np.float32)
        wob = atoms.get("wobble")
        wob = None if wob is None else np.ascontiguousarray(wob, dtype=np.float32)
        fe = np.ascontiguousarray(fe_params())
        fpp = ctypes.POINTER(ctypes.c_float)
        ctx = ctypes.c_void_p()
        _check(self._L.msct_create(ctypes.byref(p), _dp(fe), n, Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   xs[0].ctypes.data_as(fpp), xs[1].ctypes.data_as(fpp),
                                   xs[2].ctypes.data_as(fpp), occ.ctypes.data_as(fpp),
                                   None if wob is None else wob.ctypes.data_as(fpp), ctypes.byref(ctx)))
        self._ctx = ctx

    @property
    def slices(self):
        return self._L.msct_slices(self._ctx)

    @property
    def batch(self):
        """Orientations simulated together (the UTR_MSCT_BATCH_BYTES budget)."""
        return self._L.msct_batch(self._ctx)

    def images(self, angZ_deg, angY_deg, defocus, out=None):
        """[orientation][defocus][ny][nx] float32 images: intensity (output 0),
        phase (1), complex64 amplitude (2) or the 3D potential on the plane
        z = ctblength + defocus (3), as autosliccmd's noutput; angles in
        degrees about Z then Y', defocus in Angstrom.
        out: optional torch CUDA tensor of that shape (float32, complex64
        viewed as float32 pairs) filled on the device."""
        aZ = np.radians(np.ascontiguousarray(angZ_deg, dtype=np.float64))
        aY = np.radians(np.ascontiguousarray(angY_deg, dtype=np.float64))
        df = np.ascontiguousarray(np.atleast_1d(defocus), dtype=np.float64)
        if len(aZ) != len(aY):
            raise ValueError("angZ and angY differ in length")
        nx, ny, cplx_out = self.params["nx"], self.params["ny"], self.params["output"] == 2
        if out is not None:
            need = len(aZ) * len(df) * ny * nx * (2 if cplx_out else 1)
            if not (out.is_cuda and out.is_contiguous() and out.element_size() * out.numel() >= 4 * need
                    and str(out.dtype) in ("torch.float32", "torch.complex64")):
                raise ValueError("out must be a contiguous CUDA float32/complex64 tensor of at least %d floats"
                                 % need)
            _check(self._L.msct_images_device(self._ctx, len(aZ), _dp(aZ), _dp(aY), len(df), _dp(df),
                                              ctypes.c_void_p(out.data_ptr())))
            return out
        res = np.empty((len(aZ), len(df), ny, nx), np.complex64 if cplx_out else np.float32)
        _check(self._L.msct_images(self._ctx, len(aZ), _dp(aZ), _dp(aY), len(df), _dp(df),
                                   res.ctypes.data_as(ctypes.c_void_p)))
        return res

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.msct_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()


def multislice_images(params, atoms, angZ_deg, angY_deg, defocus):
    """One-shot Multislice(params, atoms).images(...)."""
    m = Multislice(params, atoms)
    try:
        return m.images(angZ_deg, angY_deg, defocus)
    finally:
        m.close()


# --------------------------------------------------------------------------
# Evaluation against a known atom set (libutr_io: pdb-compare's matching)

IO_LIB_PATH = os.path.join(_PKG, "_lib", "libutr_io.so")
_io = None


def _io_lib():
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError("libutr_io.so is not built (%s); run __graft_entry__.build()" % IO_LIB_PATH)
        L = ctypes.CDLL(IO_LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.utr_io_match_atoms.argtypes = [i, vp, vp, i, vp, vp, d, i, vp, vp, vp]
        L.utr_io_last_error.restype = ctypes.c_char_p
        _io = L
    return _io


def match_atoms(test_xyz, test_Z, ref_xyz, ref_Z, max_dist=1.0, hungarian=True):
    """pdb-compare's matching (pdb-compare.cpp:255-433) of test atoms (e.g.
    peaks of a reconstruction) to reference atoms.  Returns (match, dist,
    stats): per reference atom the matched test index and distance (-1 when
    unmatched), and a dict with matched, false_negatives, false_positives,
    mean, max and std of the matched distances."""
    L = _io_lib()
    t = np.ascontiguousarray(test_xyz, dtype=np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(ref_xyz, dtype=np.float64).reshape(-1, 3)
    tz = np.ascontiguousarray(test_Z, dtype=np.int32)
    rz = np.ascontiguousarray(ref_Z, dtype=np.int32)
    m = np.empty(len(r), np.int32)
    dist = np.empty(len(r))
    st = np.empty(6)
    rc = L.utr_io_match_atoms(len(t), t.ctypes.data, tz.ctypes.data, len(r), r.ctypes.data, rz.ctypes.data,
                              float(max_dist), 1 if hungarian else 0, m.ctypes.data, dist.ctypes.data,
                              st.ctypes.data)
    if rc:
        raise UTRError(rc, L.utr_io_last_error().decode())
    keys = ("matched", "false_negatives", "false_positives", "mean", "max", "std")
    stats = {k: (int(v) if i < 3 else float(v)) for i, (k, v) in enumerate(zip(keys, st))}
    return m, dist, stats



# --------------------------------------------------------------------------
# Learned residual corrector (utr_corr_* C ABI)

class Corrector:
    """The learned 3D residual corrector: a 3-level U-Net (BF16 on the
    matrix cores, FP32 accumulation) trained and applied on the GPU by the
    library's own kernels.  Crops are (nb, T, T, T) float32 CUDA tensors with
    T = tile + 2 * halo; volumes are (nz, ny, nx) float32 CUDA tensors."""

    def __init__(self, tile=64, halo=8, batch=4, seed=1):
        self._L = load()
        ctx = ctypes.c_void_p()
        _check(self._L.utr_corr_create(tile, halo, batch, seed, ctypes.byref(ctx)))
        self._ctx = ctx
        self.tile, self.halo, self.batch = tile, halo, batch
        self.T = tile + 2 * halo
        self.nparams = self._L.utr_corr_nparams(ctx)

    @staticmethod
    def _ptr(t, n=None):
        if not (t.is_cuda and t.is_contiguous() and str(t.dtype) == "torch.float32"):
            raise ValueError("expected a contiguous float32 CUDA tensor")
        if n is not None and t.numel() < n:
            raise ValueError("tensor holds %d floats, %d needed" % (t.numel(), n))
        return ctypes.c_void_p(t.data_ptr())

    def params(self):
        """(parameter vector float32[nparams], input scale)."""
        w = np.empty(self.nparams, np.float32)
        s = ctypes.c_float()
        _check(self._L.utr_corr_get_params(self._ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           ctypes.byref(s)))
        return w, s.value

    def set_params(self, params=None, scale=1.0):
        p = None
        if params is not None:
            params = np.ascontiguousarray(params, np.float32)
            if params.size != self.nparams:
                raise ValueError("expected %d parameters" % self.nparams)
            p = params.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        _check(self._L.utr_corr_set_params(self._ctx, p, float(scale)))

    def forward(self, crops, out):
        nb = crops.shape[0]
        n = nb * self.T ** 3
        _check(self._L.utr_corr_forward(self._ctx, self._ptr(crops, n), nb, self._ptr(out, n)))
        return out

    def gradient(self, crops, targets):
        """(loss, gradient float32[nparams]) without an update."""
        nb = crops.shape[0]
        n = nb * self.T ** 3
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_float()
        _check(self._L.utr_corr_gradient(self._ctx, self._ptr(crops, n), self._ptr(targets, n), nb,
                                         ctypes.byref(loss), g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return loss.value, g

    def train_step(self, crops, targets, lr):
        nb = crops.shape[0]
        n = nb * self.T ** 3
        loss = ctypes.c_float()
        _check(self._L.utr_corr_train_step(self._ctx, self._ptr(crops, n), self._ptr(targets, n), nb, float(lr),
                                           ctypes.byref(loss)))
        return loss.value

    def apply(self, vol, out):
        nz, ny, nx = vol.shape
        _check(self._L.utr_corr_apply(self._ctx, self._ptr(vol, vol.numel()), self._ptr(out, vol.numel()),
                                      nx, ny, nz))
        return out

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.utr_corr_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()


# --------------------------------------------------------------------------
# Evaluation against the forward model

def ground_truth(n, xst, atoms, qmax, scale=1.0, out=None):
    """Band-limited ground-truth potential of an atom set on an n^3 grid of
    step xst (utr_ground_truth_device): a float32 CUDA tensor (n, n, n),
    voxel (i, j, k) at cube position (i, j, k) * xst; atoms: dict with Z, x,
    y, z in cube coordinates (A); potential in V times scale, frequencies
    |q| < qmax (1/A)."""
    import torch
    L = load()
    Z = np.ascontiguousarray(atoms["Z"], np.int32)
    xs = [np.ascontiguousarray(atoms[k], np.float32) for k in ("x", "y", "z")]
    if out is None:
        out = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    fe = np.ascontiguousarray(fe_params())
    fpp = ctypes.POINTER(ctypes.c_float)
    _check(L.utr_ground_truth_device(n, n, n, float(xst), float(qmax), float(scale), _dp(fe), len(Z),
                                     Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), xs[0].ctypes.data_as(fpp),
                                     xs[1].ctypes.data_as(fpp), xs[2].ctypes.data_as(fpp),
                                     ctypes.c_void_p(out.data_ptr())))
    return out


def reproject(params, vol, rows, images=None, out=None):
    """Multislice reprojection of a CUDA volume (nz, ny, nx) in UTR units at
    the orientation rows (dict angZ, angY, defocus; degrees, A) with the
    error of DHT.cpp:1686 against images (CUDA tensor (rows, ny, nx) or
    None).  Returns (out or None, errors float64[rows] or None)."""
    L = load()
    p = make_params(params)
    aZ = np.ascontiguousarray(rows["angZ"], np.float64)
    aY = np.ascontiguousarray(rows["angY"], np.float64)
    df = np.ascontiguousarray(rows["defocus"], np.float64)
    err = np.zeros(len(aZ)) if images is not None else None
    for t in (vol, images, out):
        if t is not None and not (t.is_cuda and t.is_contiguous() and str(t.dtype) == "torch.float32"):
            raise ValueError("volumes and images must be contiguous float32 CUDA tensors")
    _check(L.utr_reproject_device(ctypes.byref(p), ctypes.c_void_p(vol.data_ptr()), len(aZ), _dp(aZ), _dp(aY),
                                  _dp(df), None if images is None else ctypes.c_void_p(images.data_ptr()),
                                  None if out is None else ctypes.c_void_p(out.data_ptr()),
                                  None if err is None else _dp(err)))
    return out, err
