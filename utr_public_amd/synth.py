"""Seeded synthetic UTR inputs (ctypes over _lib/libutr_synth.so).

Atom sets and images of the shape and statistics the UTR configurations
describe: analytic projections (project, images_for; flat-field corrected
intensities around 1, optional Poisson noise) for CPU-side and shape tests,
and the GPU multislice forward model's images (multislice_images_for), the
MsctKirkland output the configurations name.  Atom rotation follows
TemsimLibrary/autosliccmd.cpp:734-748.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
SYNTH_PATH = os.path.join(_PKG, "_lib", "libutr_synth.so")
_lib = None

# species of the 3000-atom configuration (C:N:O:S = 62:17:19:2)
CNOS = ([6, 7, 8, 16], [0.62, 0.17, 0.19, 0.02])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError("libutr_synth.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(SYNTH_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.utr_synth_atoms.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ip, dp, ctypes.c_int, dp, ip]
        L.utr_synth_project.argtypes = [dp, ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_int, dp, dp, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                        ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def atoms(seed, natoms, shape="sphere", size=80.0, min_sep=0.0, species=CNOS):
    """Returns (xyz[natoms,3] in UL about the volume centre, Z[natoms])."""
    L = _load()
    zs = np.ascontiguousarray(species[0], dtype=np.int32)
    fr = np.ascontiguousarray(species[1], dtype=np.float64)
    xyz = np.zeros((natoms, 3))
    Z = np.zeros(natoms, dtype=np.int32)
    rc = L.utr_synth_atoms(seed, natoms, 0 if shape == "cube" else 1, size, min_sep,
                           zs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           fr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(zs),
                           xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if rc:
        raise RuntimeError("utr_synth_atoms failed (%d)" % rc)
    return xyz, Z


def project(xyz, Z, nx, ny, xst, angZ, angY, width=0.5, contrast=0.05, dose=0.0, seed=1,
            nthreads=None):
    """Projection images [nrows, ny, nx] float32 (I = exp(-contrast*sum))."""
    L = _load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    angZ = np.ascontiguousarray(angZ, dtype=np.float64)
    angY = np.ascontiguousarray(angY, dtype=np.float64)
    n = len(angZ)
    out = np.empty((n, ny, nx), dtype=np.float32)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    rc = L.utr_synth_project(xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(Z), nx, ny, xst, n,
                             angZ.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             angY.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), width, contrast,
                             dose, seed, out.ctypes.data, nthreads)
    if rc:
        raise RuntimeError("utr_synth_project failed (%d)" % rc)
    return out


# --------------------------------------------------------------- configs

def config(name, seed=1):
    """(params, rows, atoms-spec) of a named workload.

    metric : 256^2 images, 360 rows about y (1 deg steps), 200 nm defocus,
             256^3 volume, mode 0 with Ewald correction, NSR 1/6.
    c1     : 128^3 from 36 rows about y over [0,180), 50 nm defocus.
    c2     : metric geometry with 720 rows (defocus 0 and 2000 A per angle).
    c3     : 512^3 from 720 rows about y over [0,360).
    The atom spec also carries the config's beam energy (kev) and dose
    (e/A^2) for multislice_images_for.
    """
    from . import default_params
    if name == "c1":
        p = default_params(nx=128, ny=128, xst=0.4, zlo=-25.6, zhi=25.6, wl=0.02507934,
                           escc=1, nsr=0.0, ctf_mode=0)
        angY = np.arange(36) * 5.0
        rows = dict(angZ=np.zeros(36), angY=angY, defocus=np.full(36, 500.0))
        at = dict(natoms=400, shape="cube", size=40.0, min_sep=1.4, species=([6], [1.0]), kev=200.0,
                  dose=0.0)
    elif name in ("metric", "c2"):
        p = default_params(nx=256, ny=256, xst=0.6, zlo=-76.8, zhi=76.8, wl=0.01968749,
                           escc=1, nsr=1.0 / 6.0, ctf_mode=0)
        if name == "metric":
            angY = np.arange(360) * 1.0
            rows = dict(angZ=np.zeros(360), angY=angY, defocus=np.full(360, 2000.0))
        else:
            angY = np.repeat(np.arange(360) * 1.0, 2)
            df = np.tile([0.0, 2000.0], 360)
            rows = dict(angZ=np.zeros(720), angY=angY, defocus=df)
        at = dict(natoms=3000, shape="sphere", size=80.0, min_sep=0.0, species=CNOS, kev=300.0,
                  dose=100.0)
    elif name == "c3":
        p = default_params(nx=512, ny=512, xst=0.6, zlo=-153.6, zhi=153.6, wl=0.01968749,
                           escc=1, nsr=1.0 / 6.0, ctf_mode=0)
        angY = np.arange(720) * 0.5
        rows = dict(angZ=np.zeros(720), angY=angY, defocus=np.full(720, 2000.0))
        at = dict(natoms=12000, shape="sphere", size=160.0, min_sep=0.0, species=CNOS, kev=300.0,
                  dose=100.0)
    else:
        raise KeyError(name)
    return p, rows, at


def c4_jobs(seed=1):
    """The batched workload: 16 independent reconstructions with cube edges
    drawn from {64, 96, 128, 192, 256, 384}, angle counts from {30, 90, 180,
    360} about y, 1-3 defocus distances in [0, 300] nm per angle (flattened to
    rows), random C/N/O/S atoms at 0.05 atoms/nm^3.  Returns a list of
    (params, rows, atoms-spec)."""
    from . import default_params
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(16):
        n = int(rng.choice([64, 96, 128, 192, 256, 384]))
        na = int(rng.choice([30, 90, 180, 360]))
        nd = int(rng.integers(1, 4))
        xst = 0.6
        half = n * xst / 2
        p = default_params(nx=n, ny=n, xst=xst, zlo=-half, zhi=half, wl=0.01968749, escc=1,
                           nsr=1.0 / 6.0, ctf_mode=0)
        ang = np.arange(na) * (360.0 / na)
        df = np.sort(rng.uniform(0.0, 3000.0, nd))
        rows = dict(angZ=np.zeros(na * nd), angY=np.repeat(ang, nd), defocus=np.tile(df, na))
        side_nm = n * xst / 10.0 * 0.8
        natoms = max(8, int(0.05 * side_nm ** 3))
        at = dict(natoms=natoms, shape="cube", size=n * xst * 0.8, min_sep=0.0, species=CNOS)
        jobs.append((p, rows, at))
    return jobs


def multislice_images_for(params, rows, at, seed=1, dose=None, kev=None, deltaz=1.0):
    """Images of a config from the GPU multislice forward model (msct_*): the
    config's seeded atom set inside the params' cube (nx * xst per side),
    imaged at every row's orientation and defocus, then Poisson counts at
    `dose` electrons per A^2 (default: the config's, none if it has none).
    Atoms farther than 0.49 of the cube side from the centre are left out,
    so that every orientation keeps them inside the cube (the multislice
    stops on an atom outside it, as autosliccmd does).  Returns float32
    [rows, ny, nx] in row order.  Needs a GPU."""
    import torch
    from . import Multislice, default_msct_params
    nx, ny, xst = params["nx"], params["ny"], params["xst"]
    if nx != ny:
        raise ValueError("multislice images need a square image")
    L = nx * xst
    if kev is None:
        kev = at.get("kev")
    if kev is None:
        # invert wl = 12.3984244 / sqrt(kev (2 m0c2 + kev)) (slicelib.cpp:2891-2900)
        m0c2 = 510.99906
        kev = -m0c2 + (m0c2 * m0c2 + (12.3984244 / params["wl"]) ** 2) ** 0.5
    if dose is None:
        dose = at.get("dose", 0.0)
    xyz, Z = atoms(seed, at["natoms"], at["shape"], at["size"], at["min_sep"], at["species"])
    keep = np.sqrt((xyz ** 2).sum(1)) <= 0.49 * L
    xyz, Z = xyz[keep], Z[keep]
    h = L / 2
    spec = dict(Z=Z, x=(xyz[:, 0] + h).astype(np.float32), y=(xyz[:, 1] + h).astype(np.float32),
                z=(xyz[:, 2] + h).astype(np.float32), occ=np.ones(len(Z), np.float32))
    mp = default_msct_params(nx=nx, ny=ny, ctblength=L, deltaz=deltaz, kev=float(kev))
    angZ = np.asarray(rows["angZ"], np.float64)
    angY = np.asarray(rows["angY"], np.float64)
    df = np.asarray(rows["defocus"], np.float64)
    out = torch.empty((len(df), ny, nx), dtype=torch.float32, device="cuda")
    m = Multislice(mp, spec)
    try:
        for d in np.unique(df):
            idx = np.nonzero(df == d)[0]
            part = torch.empty((len(idx), 1, ny, nx), dtype=torch.float32, device="cuda")
            m.images(angZ[idx], angY[idx], [float(d)], out=part)
            out[torch.from_numpy(idx).cuda()] = part[:, 0]
    finally:
        m.close()
    if dose > 0:
        counts = dose * xst * xst
        g = torch.Generator(device="cuda")
        g.manual_seed(int(seed))
        out = torch.poisson(torch.clamp(out, min=0.0) * counts, generator=g) / counts
    torch.cuda.synchronize()
    return out.cpu().numpy()


def images_for(params, rows, at, seed=1, dose=None):
    """Synthetic images for a config: dose defaults to (1/nsr)^2 counts/px."""
    xyz, Z = atoms(seed, at["natoms"], at["shape"], at["size"], at["min_sep"], at["species"])
    if dose is None:
        dose = (1.0 / params["nsr"]) ** 2 if params["nsr"] > 0 else 0.0
    return project(xyz, Z, params["nx"], params["ny"], params["xst"], rows["angZ"], rows["angY"],
                   dose=dose, seed=seed)
