"""Python mirror of libwiremom's C ABI (include/wiremom.h).

The compute path is the HIP library ``libwiremom.so`` built in this directory
for gfx950; this module only marshals arguments.  There is no CPU fallback:
if the library is missing, every entry point raises.

Reference interfaces mirrored here (same names, argument meaning and
units): ``geoinp``/``configwire`` ("wersja referencyjna/Configuration.cpp":5-75,
112-520), ``zmatrix`` ("wersja referencyjna/Zmatrix.cpp":27-52),
``zmatrix_gpu`` ("wersja referencyjna/Wire_MoM.cu":4052-4197), ``LU_cuBLAS``
and ``CalculateByFreq`` ("wersja gpu/Wire_MoM.cu":355-504).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwiremom.so")

PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
PREC_NAMES = {"f64": PREC_F64, "f32": PREC_F32, "bf16x3": PREC_BF16X3}
# status codes (WM_ERR_*)
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_SPIN, ERR_NOCONV, ERR_IO, ERR_GEOMETRY, ERR_STATE = -1, -2, -3, -4, -5, -6, -7, -8
# capacities of the reference driver ("wersja referencyjna/Wire_MoM.cu":43-44)
NNMAX, NTMAX, NWMAX, NPWMAX = 3000, 5000, 10000, 20480
NSMAX = NWMAX + NPWMAX
# zmatrix's environment signs (sgnx, sgny, sgnz, sgnenv) for ienv 1 (free
# space) and 2 (PEC ground z = 0), "wersja referencyjna/Wire_MoM.cu":107-122
ENV_SIGNS = {1: (1.0, 1.0, 1.0, 1.0), 2: (1.0, 1.0, -1.0, -1.0)}

EXPORTS = [
    "wm_last_error", "wm_version", "wm_init", "wm_shutdown", "wm_sync", "wm_device_info", "wm_prof_enable",
    "wm_prof_get", "wm_dev_alloc",
    "wm_dev_free", "wm_memcpy_h2d", "wm_memcpy_d2h", "wm_memcpy_d2d", "wm_geom_create", "wm_geom_create_env", "wm_geom_load", "wm_geom_load_env", "wm_geom_destroy",
    "wm_geom_npls", "wm_geom_ld", "wm_zfill", "wm_getrf_batched", "wm_getrs_batched", "wm_set_getrf_block", "wm_get_getrf_block", "wm_tune_getrf", "wm_set_fill_freqs", "wm_get_fill_freqs",
    "wm_tune_fill", "wm_tune_save", "wm_tune_load", "wm_set_graphs", "wm_set_fill_overlap", "wm_solve_opts_default",
    "wm_solve_sweep", "wm_solve_sweep_ex", "wm_transient", "wm_spectrum_dexp", "wm_server_opts_default", "wm_server_create", "wm_server_submit", "wm_server_step",
    "wm_server_result", "wm_server_status", "wm_server_done", "wm_server_destroy", "geoinp", "config", "configwire", "tablice", "zmatrix",
    "zmatrix_gpu", "zmatrix_gpu_rank", "LU_cuBLAS", "CalculateByFreq",
]


class WireMoMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libwiremom error {code}: {msg}")
        self.code = code


class SolveOpts(C.Structure):
    _fields_ = [("prec", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double), ("batch", C.c_int),
                ("hbm_fraction", C.c_double), ("fallback_f64", C.c_int), ("size_policy", C.c_int),
                ("cond_est", C.c_int), ("recycle", C.c_double)]


def _prec(p) -> int:
    return PREC_NAMES[p] if isinstance(p, str) else int(p)


def _stats(s) -> dict:
    return dict(iters=s.iters, info=s.info, fell_back=s.fell_back, prec=s.prec, rate=s.rate, dx_rel=s.dx_rel,
                cond1=s.cond1, anorm1=s.anorm1, anchor_mhz=s.anchor_mhz, berr=s.berr)


class FreqStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("info", C.c_int), ("fell_back", C.c_int), ("prec", C.c_int),
                ("rate", C.c_double), ("dx_rel", C.c_double), ("cond1", C.c_double), ("anorm1", C.c_double),
                ("anchor_mhz", C.c_double), ("berr", C.c_double)]


class Load(C.Structure):
    """wm_load: series load r + j(2 pi f l - 1/(2 pi f c)) on a basis (f MHz, l uH, c uF)."""
    _fields_ = [("basis", C.c_int), ("r_ohm", C.c_double), ("l_uh", C.c_double), ("c_uf", C.c_double)]


class Source(C.Structure):
    """wm_source: delta-gap generator ampl*exp(j pi phas/180) on a basis, in right-hand side rhs."""
    _fields_ = [("basis", C.c_int), ("rhs", C.c_int), ("ampl", C.c_float), ("phas", C.c_float)]


class C128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class PlaneWave(C.Structure):
    """wm_plane_wave: arrives from (theta, phi) degrees, field e_theta th + e_phi ph, in right-hand side rhs."""
    _fields_ = [("rhs", C.c_int), ("theta_deg", C.c_double), ("phi_deg", C.c_double), ("e_theta", C128),
                ("e_phi", C128)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_inflight", C.c_int), ("prec", C.c_int), ("streams", C.c_int),
                ("size_policy", C.c_int)]


_lib = None


def _p(t):
    return C.POINTER(t)


def lib() -> C.CDLL:
    """Loads libwiremom.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python praca-magisterska_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, ip, fp, dp = C.c_void_p, _p(C.c_int), _p(C.c_float), _p(C.c_double)
        L.wm_last_error.restype = C.c_char_p
        L.wm_version.restype = C.c_char_p
        L.wm_init.argtypes = [C.c_int, _p(vp)]
        L.wm_shutdown.argtypes = [vp]
        L.wm_sync.argtypes = [vp]
        L.wm_prof_enable.argtypes = [vp, C.c_int]
        L.wm_prof_get.argtypes = [vp, C.c_char_p, dp, _p(C.c_int64), dp, dp]
        L.wm_device_info.argtypes = [vp, C.c_char_p, C.c_int, ip, _p(C.c_size_t)]
        L.wm_dev_alloc.argtypes = [vp, C.c_size_t, _p(vp)]
        L.wm_dev_free.argtypes = [vp, vp]
        L.wm_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
        L.wm_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
        L.wm_memcpy_d2d.argtypes = [vp, vp, vp, C.c_size_t]
        L.wm_geom_create.argtypes = [vp, fp, fp, fp, fp, ip, ip, ip, C.c_int, C.c_int, _p(vp)]
        L.wm_geom_load.argtypes = [vp, C.c_char_p, C.c_char_p, _p(vp)]
        L.wm_geom_create_env.argtypes = [vp, fp, fp, fp, fp, ip, ip, ip, C.c_int, C.c_int, C.c_int, C.c_float,
                                         C.c_float, C.c_float, C.c_float, _p(vp)]
        L.wm_geom_load_env.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int, _p(vp)]
        L.wm_geom_destroy.argtypes = [vp]
        L.wm_geom_npls.argtypes = [vp]
        L.wm_geom_ld.argtypes = [vp]
        L.wm_zfill.argtypes = [vp, vp, C.c_int, dp, vp, C.c_int64, vp, C.c_int64]
        L.wm_getrf_batched.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, C.c_int64, vp, C.c_int64, C.c_int, ip]
        L.wm_getrs_batched.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int64, vp, C.c_int64, vp,
                                       C.c_int, C.c_int64, C.c_int]
        L.wm_set_getrf_block.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.wm_set_graphs.argtypes = [vp, C.c_int]
        L.wm_set_fill_overlap.argtypes = [vp, C.c_int]
        L.wm_get_getrf_block.argtypes = [vp, C.c_int, C.c_int]
        L.wm_tune_getrf.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, dp]
        L.wm_set_fill_freqs.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.wm_get_fill_freqs.argtypes = [vp, C.c_int, C.c_int]
        L.wm_tune_fill.argtypes = [vp, vp, ip, dp]
        L.wm_tune_save.argtypes = [vp, C.c_char_p]
        L.wm_tune_load.argtypes = [vp, C.c_char_p]
        L.wm_solve_opts_default.argtypes = [_p(SolveOpts)]
        L.wm_solve_opts_default.restype = None
        L.wm_solve_sweep.argtypes = [vp, vp, _p(SolveOpts), C.c_int, fp, C.c_int, ip, C.c_float, C.c_float, vp, vp,
                                     _p(FreqStats)]
        L.wm_solve_sweep_ex.argtypes = [vp, vp, _p(SolveOpts), C.c_int, fp, C.c_int, C.c_int, _p(Source), C.c_int,
                                        _p(PlaneWave), C.c_int, _p(Load), vp, vp, _p(FreqStats)]
        L.wm_transient.argtypes = [vp, C.c_int, fp, C.c_int, vp, vp, C.c_int, C.c_double, C.c_double, vp]
        L.wm_spectrum_dexp.argtypes = [C.c_int, fp, C.c_double, C.c_double, C.c_double, vp]
        L.wm_server_opts_default.argtypes = [_p(ServerOpts)]
        L.wm_server_opts_default.restype = None
        L.wm_server_create.argtypes = [vp, _p(ServerOpts), _p(vp)]
        L.wm_server_submit.argtypes = [vp, vp, C.c_int, fp, C.c_int, ip, C.c_float, C.c_float, _p(C.c_int64)]
        L.wm_server_step.argtypes = [vp, C.c_int, ip]
        L.wm_server_result.argtypes = [vp, C.c_int64, vp, vp, _p(FreqStats), dp]
        L.wm_server_status.argtypes = [vp, ip, ip, ip, _p(C.c_size_t)]
        L.wm_server_done.argtypes = [vp, _p(C.c_int64), C.c_int, ip]
        L.wm_server_destroy.argtypes = [vp]
        L.wm_server_overlap.argtypes = [vp, _p(C.c_int64), dp, dp]
        L.geoinp.argtypes = [C.c_int, C.c_int, C.c_int, ip, fp, fp, fp, ip, ip, fp, ip, ip, ip]
        L.geoinp.restype = None
        L.configwire.argtypes = [ip, C.c_int, fp, fp, fp, fp, fp, fp, C.c_int, ip, fp, ip, ip, ip, ip, ip, ip, fp,
                                 ip, ip, C.c_int, C.c_int, C.c_int]
        L.configwire.restype = None
        L.zmatrix.argtypes = [vp, fp, fp, fp, ip, ip, ip, fp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_float,
                              C.c_float, C.c_float, C.c_float]
        L.zmatrix.restype = None
        L.zmatrix_gpu.argtypes = [vp, fp, fp, fp, fp, ip, ip, ip, C.c_int, C.c_double, C.c_int, C.c_int, C.c_float,
                                  C.c_float, C.c_float, C.c_float]
        L.zmatrix_gpu.restype = None
        L.LU_cuBLAS.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_int]
        L.LU_cuBLAS.restype = None
        L.CalculateByFreq.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, fp, fp, fp, fp, ip, ip, ip, C.c_int,
                                      C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_float,
                                      C.c_int, C.c_float, vp, vp, vp, C.c_int]
        L.CalculateByFreq.restype = None
        _lib = L
    return _lib


def _check(rc: int) -> int:
    if rc < 0:
        raise WireMoMError(rc, lib().wm_last_error().decode(errors="replace"))
    return rc


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(_p(t))


def wavenumber(freq_mhz: float) -> float:
    """k for a frequency in MHz stored as float32 ("wersja referencyjna/Wire_MoM.cu":5-6,338-340)."""
    f = float(np.float32(freq_mhz))
    ambda = 2.99792448E+8 / f * 1E-6
    return (np.pi + np.pi) / ambda


# ---------------------------------------------------------------------------- front end
@dataclass
class Topology:
    """Arrays after configwire (0-based positions, 1-based ids, stride nsmax)."""
    nsmax: int
    npnt: int
    nseg: int
    npls: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    rad: np.ndarray
    ise: np.ndarray
    isp: np.ndarray
    ips: np.ndarray
    ipc: np.ndarray
    iwe: np.ndarray


def read_geo(path: str):
    """Parses a .geo file into 1-based tables (same reading rules as geoinp)."""
    with open(path, "rb") as f:
        f.readline()
        f.readline()
        tok = f.read().split()
    it = iter(tok)
    nnod = int(next(it))
    xn = np.zeros(nnod + 1, np.float32)
    yn = np.zeros_like(xn)
    zn = np.zeros_like(xn)
    for i in range(1, nnod + 1):
        xn[i], yn[i], zn[i] = (np.float32(float(next(it))) for _ in range(3))
    nw = int(next(it))
    iwn = np.zeros(2 * nw + 1, np.int32)
    wrad = np.zeros(nw + 1, np.float32)
    nsw = np.zeros(nw + 1, np.int32)
    for i in range(1, nw + 1):
        iwn[i] = int(next(it))
        iwn[i + nw] = int(next(it))
        wrad[i] = np.float32(float(next(it)))
        nsw[i] = int(next(it))
    return nnod, xn, yn, zn, nw, iwn, wrad, nsw


def configwire(geo_path: str, ienv: int = 1, nsmax: int = NSMAX, nwmax: int = NWMAX, cwd: str | None = None) -> Topology:
    """Runs the library's configwire (reference parameter list) on a .geo file.

    configwire.txt is written to ``cwd`` (default: the current directory)."""
    nnod, xn, yn, zn, nw, iwn, wrad, nsw = read_geo(geo_path)
    x = np.zeros(nsmax, np.float32)
    y = np.zeros_like(x)
    z = np.zeros_like(x)
    rad = np.zeros_like(x)
    ise = np.zeros(2 * nsmax, np.int32)
    isp = np.zeros_like(ise)
    ips = np.zeros_like(ise)
    iwe = np.zeros_like(ise)
    ipc = np.zeros(nsmax, np.int32)
    npnt, npls, nseg = C.c_int(), C.c_int(), C.c_int()
    old = os.getcwd()
    if cwd:
        os.chdir(cwd)
    try:
        lib().configwire(C.byref(npnt), nnod, _ptr(xn, C.c_float), _ptr(yn, C.c_float), _ptr(zn, C.c_float),
                         _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(z, C.c_float), nw, _ptr(iwn, C.c_int),
                         _ptr(wrad, C.c_float), _ptr(nsw, C.c_int), _ptr(ise, C.c_int), _ptr(isp, C.c_int),
                         _ptr(ips, C.c_int), _ptr(ipc, C.c_int), _ptr(iwe, C.c_int), _ptr(rad, C.c_float),
                         C.byref(npls), C.byref(nseg), nsmax, nwmax, ienv)
    finally:
        os.chdir(old)
    return Topology(nsmax, npnt.value, nseg.value, npls.value, x, y, z, rad, ise, isp, ips, ipc, iwe)


# ---------------------------------------------------------------------------- device objects
class Context:
    """A device context (wm_init)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().wm_init(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().wm_shutdown(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(lib().wm_sync(self.h))

    def prof_enable(self, on: bool = True):
        _check(lib().wm_prof_enable(self.h, int(on)))

    def prof_get(self, kernel: str):
        """(ms, launches, algorithmic flops, algorithmic bytes) of one kernel class."""
        ms, fl, by = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int64()
        _check(lib().wm_prof_get(self.h, kernel.encode(), C.byref(ms), C.byref(n), C.byref(fl), C.byref(by)))
        return ms.value, n.value, fl.value, by.value

    def set_graphs(self, on: bool):
        """Replay factorizations from captured HIP graphs (default) or launch directly."""
        _check(lib().wm_set_graphs(self.h, int(on)))

    def set_fill_overlap(self, on: bool):
        """Fill the GMRES frequencies of recycled sweeps on a second stream while the anchors are factored
        (default) or keep every launch on one stream."""
        _check(lib().wm_set_fill_overlap(self.h, int(on)))

    def set_getrf_block(self, prec, n: int, panels: int):
        """Panels per trailing update of the factorization for (prec, size class of n); 0 = default."""
        _check(lib().wm_set_getrf_block(self.h, _prec(prec), n, panels))

    def getrf_block(self, prec, n: int) -> int:
        return lib().wm_get_getrf_block(self.h, _prec(prec), n)

    def set_fill_freqs(self, ienv: int, n: int, freqs: int):
        """Frequencies per fill launch (1-4, 0 = default) for environment ienv and the size class of n."""
        _check(lib().wm_set_fill_freqs(self.h, ienv, n, freqs))

    def fill_freqs(self, ienv: int, n: int) -> int:
        return lib().wm_get_fill_freqs(self.h, ienv, n)

    def tune_fill(self, geom):
        """Times the fill of geom at 1-4 frequencies per launch, keeps the fastest; returns (freqs, ms per frequency)."""
        best, ms = C.c_int(), C.c_double()
        _check(lib().wm_tune_fill(self.h, geom.h, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def tune_save(self, path: str) -> int:
        """Writes this context's tuning settings as JSON lines; returns the number of entries."""
        n = lib().wm_tune_save(self.h, path.encode())
        if n < 0:
            _check(n)
        return n

    def tune_load(self, path: str) -> int:
        """Applies the entries of a tuning cache made on this architecture; returns how many."""
        n = lib().wm_tune_load(self.h, path.encode())
        if n < 0:
            _check(n)
        return n

    def tune_getrf(self, prec, n: int, batch: int = 0):
        """Times each outer-block depth on synthetic matrices, keeps the fastest; returns (panels, ms)."""
        best, ms = C.c_int(), C.c_double()
        _check(lib().wm_tune_getrf(self.h, _prec(prec), n, batch, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        hbm = C.c_size_t()
        _check(lib().wm_device_info(self.h, name, 256, C.byref(cu), C.byref(hbm)))
        return name.value.decode(), cu.value, hbm.value


class DeviceArray:
    """Device buffer holding a copy of a numpy array (for the batched LU entry points)."""

    def __init__(self, ctx: Context, arr: np.ndarray):
        self.ctx = ctx
        self.shape, self.dtype = arr.shape, arr.dtype
        self.nbytes = arr.nbytes
        p = C.c_void_p()
        _check(lib().wm_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p
        self.upload(arr)

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        _check(lib().wm_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(lib().wm_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().wm_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Geometry:
    """Device-resident geometry (wm_geom_create / wm_geom_load)."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.h = ctx, handle
        self.npls = lib().wm_geom_npls(handle)
        self.ld = lib().wm_geom_ld(handle)

    @classmethod
    def from_topology(cls, ctx: Context, t: Topology, ienv: int = 1) -> "Geometry":
        """ienv 1 free space, 2 PEC ground z = 0 (the reference's image signs)."""
        h = C.c_void_p()
        sx, sy, sz, se = ENV_SIGNS[ienv] if ienv in ENV_SIGNS else (1.0, 1.0, 1.0, 1.0)
        _check(lib().wm_geom_create_env(ctx.h, _ptr(t.x, C.c_float), _ptr(t.y, C.c_float), _ptr(t.z, C.c_float),
                                        _ptr(t.rad, C.c_float), _ptr(t.ipc, C.c_int), _ptr(t.ips, C.c_int),
                                        _ptr(t.ise, C.c_int), t.nsmax, t.npls, int(ienv), sx, sy, sz, se,
                                        C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx: Context, geo_path: str, wirefile: str | None = None, ienv: int = 1) -> "Geometry":
        """Reads a .geo and segments it (configwire) for environment ienv."""
        h = C.c_void_p()
        _check(lib().wm_geom_load_env(ctx.h, geo_path.encode(), (wirefile or "").encode(), int(ienv), C.byref(h)))
        return cls(ctx, h)

    def close(self):
        if self.h:
            lib().wm_geom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fill(self, freqs_mhz, f32_copy: bool = False):
        """Z matrices (complex128, N x N each) for the given frequencies, filled on the GPU."""
        ka = np.array([wavenumber(f) for f in np.atleast_1d(freqs_mhz)], np.float64)
        nf, ld, n = len(ka), self.ld, self.npls
        Zd = DeviceArray(self.ctx, np.zeros((nf, ld, ld), np.complex128))
        Z32 = DeviceArray(self.ctx, np.zeros((nf, ld, ld), np.complex64)) if f32_copy else None
        _check(lib().wm_zfill(self.ctx.h, self.h, nf, _ptr(ka, C.c_double), Zd.ptr, ld * ld,
                              Z32.ptr if Z32 else None, ld * ld))
        # device layout is column-major: element (i, j) at j*ld + i
        Z = Zd.download().transpose(0, 2, 1)
        out = [Z[:, :n, :n]]
        if Z32:
            out.append(Z32.download().transpose(0, 2, 1)[:, :n, :n])
        return out if f32_copy else out[0]


def solve_sweep(geom: Geometry, freqs_mhz, gaps, prec="f64", ampl: float = 1.0, phas: float = 0.0,
                batch: int = 0, max_iter: int = 20, tol: float = 1e-12, fallback: bool = True,
                size_policy: bool = True, cond_est: bool = False, return_status: bool = False,
                recycle: float = 0.0):
    """Whole solves (wm_solve_sweep).  Returns (zin[nf, nrhs], I[nf, nrhs, N], stats list).
    size_policy=False keeps the requested mixed precision at sizes where FP64 factors are cheaper;
    cond_est=True fills stats[f]["cond1"] with an estimate of the 1-norm condition number of Z.
    return_status=True appends the call's status to the result instead of raising on
    ERR_NOCONV (refinement that did not converge with fallback=False).
    recycle=r > 1 factors only anchor frequencies (every frequency within a factor r of one) and
    solves the others by FP64 GMRES refinement with the nearest anchor's factors."""
    f = np.ascontiguousarray(np.atleast_1d(freqs_mhz), np.float32)
    g = np.ascontiguousarray(np.atleast_1d(gaps), np.int32)
    nf, nr, n = len(f), len(g), geom.npls
    o = SolveOpts()
    lib().wm_solve_opts_default(C.byref(o))
    o.prec = PREC_NAMES[prec] if isinstance(prec, str) else int(prec)
    o.batch, o.max_iter, o.tol, o.fallback_f64 = batch, max_iter, tol, int(fallback)
    o.size_policy = int(size_policy)
    o.cond_est = int(cond_est)
    o.recycle = float(recycle)
    I = np.zeros((nf, nr, n), np.complex128)
    zin = np.zeros((nf, nr), np.complex128)
    st = (FreqStats * nf)()
    rc = lib().wm_solve_sweep(geom.ctx.h, geom.h, C.byref(o), nf, _ptr(f, C.c_float), nr, _ptr(g, C.c_int),
                              float(ampl), float(phas), I.ctypes.data, zin.ctypes.data, st)
    if not (return_status and rc == ERR_NOCONV):
        _check(rc)
    stats = [_stats(s) for s in st]
    return (zin, I, stats, rc) if return_status else (zin, I, stats)


def solve_sweep_ex(geom: Geometry, freqs_mhz, nrhs: int, sources, loads=(), prec="f64", batch: int = 0,
                   max_iter: int = 20, tol: float = 1e-12, fallback: bool = True, size_policy: bool = True,
                   plane_waves=(), recycle: float = 0.0, cond_est: bool = False):
    """Sweep with general generators, plane waves and lumped loads (wm_solve_sweep_ex).

    sources: (basis, rhs, ampl, phas) per generator; loads: (basis, r_ohm, l_uh, c_uf) per load;
    plane_waves: (rhs, theta_deg, phi_deg, e_theta, e_phi) per wave (complex field components).
    Returns (zin[nf, nsrc] = V/I at each generator, I[nf, nrhs, N], stats list)."""
    f = np.ascontiguousarray(np.atleast_1d(freqs_mhz), np.float32)
    nf, n = len(f), geom.npls
    src = (Source * max(1, len(sources)))(*[Source(int(b), int(r), float(a), float(p)) for b, r, a, p in sources])
    lds = (Load * max(1, len(loads)))(*[Load(int(b), float(r), float(l), float(c)) for b, r, l, c in loads])
    pws = (PlaneWave * max(1, len(plane_waves)))(
        *[PlaneWave(int(r), float(t), float(p), C128(complex(et).real, complex(et).imag),
                    C128(complex(ep).real, complex(ep).imag)) for r, t, p, et, ep in plane_waves])
    o = SolveOpts()
    lib().wm_solve_opts_default(C.byref(o))
    o.prec = PREC_NAMES[prec] if isinstance(prec, str) else int(prec)
    o.batch, o.max_iter, o.tol, o.fallback_f64 = batch, max_iter, tol, int(fallback)
    o.size_policy = int(size_policy)
    o.recycle, o.cond_est = float(recycle), int(cond_est)
    I = np.zeros((nf, nrhs, n), np.complex128)
    zin = np.zeros((nf, len(sources)), np.complex128)
    st = (FreqStats * nf)()
    rc = lib().wm_solve_sweep_ex(geom.ctx.h, geom.h, C.byref(o), nf, _ptr(f, C.c_float), int(nrhs), len(sources),
                                 src, len(plane_waves), pws, len(loads), lds, I.ctypes.data,
                                 zin.ctypes.data if len(sources) else None, st)
    _check(rc)
    return zin, I, [_stats(s) for s in st]


def spectrum_dexp(freqs_mhz, i0: float, alpha: float, beta: float) -> np.ndarray:
    """Spectrum of i0 (e^{-alpha t} - e^{-beta t}) at the given frequencies (wm_spectrum_dexp)."""
    f = np.ascontiguousarray(np.atleast_1d(freqs_mhz), np.float32)
    S = np.zeros(len(f), np.complex128)
    _check(lib().wm_spectrum_dexp(len(f), _ptr(f, C.c_float), float(i0), float(alpha), float(beta), S.ctypes.data))
    return S


def transient(ctx: Context, freqs_mhz, H, S=None, t0: float = 0.0, dt: float = 1e-7, nt: int = 1024) -> np.ndarray:
    """Real time responses r[m, j] at t0 + j dt from transfers H[k, m] (frequency-major) and spectrum S[k]
    (wm_transient)."""
    f = np.ascontiguousarray(np.atleast_1d(freqs_mhz), np.float32)
    H = np.ascontiguousarray(np.asarray(H, np.complex128).reshape(len(f), -1))
    Sv = None if S is None else np.ascontiguousarray(np.asarray(S, np.complex128))
    r = np.zeros((H.shape[1], nt), np.float64)
    _check(lib().wm_transient(ctx.h, len(f), _ptr(f, C.c_float), H.shape[1], H.ctypes.data,
                              None if Sv is None else Sv.ctypes.data, int(nt), float(t0), float(dt), r.ctypes.data))
    return r


def getrf_batched(ctx: Context, A: np.ndarray, prec="f64"):
    """LU of a stack of n x n matrices (n a multiple of 64) on the GPU.

    Returns (LU stack as numpy, ipiv 1-based, info list); complex64 for mixed precisions."""
    p = PREC_NAMES[prec] if isinstance(prec, str) else int(prec)
    dt = np.complex128 if p == PREC_F64 else np.complex64
    A = np.asarray(A)
    if A.ndim == 2:
        A = A[None]
    b, n, _ = A.shape
    dev = DeviceArray(ctx, np.ascontiguousarray(A.transpose(0, 2, 1).astype(dt)))  # column-major
    piv = DeviceArray(ctx, np.zeros((b, n), np.int32))
    info = np.zeros(b, np.int32)
    rc = lib().wm_getrf_batched(ctx.h, p, n, dev.ptr, n, n * n, piv.ptr, n, b, _ptr(info, C.c_int))
    _check(rc)
    return dev, piv, info


def getrs_batched(ctx: Context, dev: DeviceArray, piv: DeviceArray, B: np.ndarray, prec="f64") -> np.ndarray:
    p = PREC_NAMES[prec] if isinstance(prec, str) else int(prec)
    b, n = piv.shape
    B = np.asarray(B, np.complex128)
    if B.ndim == 2:
        B = B[None]
    nr = B.shape[2]
    Bd = DeviceArray(ctx, np.ascontiguousarray(B.transpose(0, 2, 1)))
    _check(lib().wm_getrs_batched(ctx.h, p, n, nr, dev.ptr, n, n * n, piv.ptr, n, Bd.ptr, n, n * nr, b))
    return Bd.download().transpose(0, 2, 1)


# ---------------------------------------------------------------------------- drop-in wrappers
def zmatrix(t: Topology, ka: float, ienv: int = 1) -> np.ndarray:
    """zmatrix with the reference's arguments; returns Z (npls x npls, numpy row/col indexing)."""
    n = t.npls
    cz = np.zeros((n, n), np.complex128)  # filled column-major: cz.T is Z
    sx, sy, sz, se = ENV_SIGNS[ienv]
    lib().zmatrix(cz.ctypes.data, _ptr(t.x, C.c_float), _ptr(t.y, C.c_float), _ptr(t.z, C.c_float),
                  _ptr(t.ipc, C.c_int), _ptr(t.ips, C.c_int), _ptr(t.ise, C.c_int), _ptr(t.rad, C.c_float), n,
                  float(ka), t.nsmax, int(ienv), sx, sy, sz, se)
    return cz.T.copy()


def LU_cuBLAS(Z: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Solves Z x = b with the drop-in LU_cuBLAS (FP64 on the GPU)."""
    n = Z.shape[0]
    cz = np.ascontiguousarray(np.asarray(Z, np.complex128).T)  # column-major
    B = np.asarray(b, np.complex128)
    nr = 1 if B.ndim == 1 else B.shape[1]
    # ci is the in/out buffer of the reference contract: always a fresh copy,
    # never a view of the caller's array
    ci = np.array(B.reshape(n, nr).T, dtype=np.complex128, order="C", copy=True)
    cu = np.zeros_like(ci)
    lib().LU_cuBLAS(n, nr, ci.ctypes.data, cu.ctypes.data, cz.ctypes.data, 1, 0)
    x = ci.T
    return x[:, 0] if B.ndim == 1 else x


def CalculateByFreq(t: Topology, freq_mhz: float, ngap: int, ampl: float = 1.0, phas: float = 0.0,
                    iload: int = 2, rank: int = 0, ienv: int = 1):
    """One frequency through the drop-in CalculateByFreq; returns (Zin, I, Z)."""
    n = t.npls
    sx, sy, sz, se = ENV_SIGNS[ienv]
    cz = np.zeros((n, n), np.complex128)
    cu = np.zeros(n, np.complex128)
    ci = np.zeros(n, np.complex128)
    lib().CalculateByFreq(0, n, 1, float(freq_mhz), _ptr(t.x, C.c_float), _ptr(t.y, C.c_float),
                          _ptr(t.z, C.c_float), _ptr(t.rad, C.c_float), _ptr(t.ipc, C.c_int), _ptr(t.ips, C.c_int),
                          _ptr(t.ise, C.c_int), t.nsmax, int(ienv), sx, sy, sz, se, iload, float(phas), int(ngap),
                          float(ampl), cz.ctypes.data, cu.ctypes.data, ci.ctypes.data, rank)
    return cu[ngap - 1] / ci[ngap - 1], ci, cz.T.copy()


class Server:
    """Continuous-batching request server (wm_server_*)."""

    def __init__(self, ctx: Context, prec="f32", hbm_budget: int = 0, max_inflight: int = 0,
                 size_policy: bool = True, streams: int = 2):
        o = ServerOpts()
        lib().wm_server_opts_default(C.byref(o))
        o.prec = PREC_NAMES[prec] if isinstance(prec, str) else int(prec)
        o.hbm_budget, o.max_inflight, o.size_policy = hbm_budget, max_inflight, int(size_policy)
        o.streams = int(streams)
        h = C.c_void_p()
        _check(lib().wm_server_create(ctx.h, C.byref(o), C.byref(h)))
        self.h, self.ctx = h, ctx
        self.meta = {}

    def submit(self, geom: Geometry, freqs_mhz, gaps, ampl: float = 1.0, phas: float = 0.0) -> int:
        f = np.ascontiguousarray(np.atleast_1d(freqs_mhz), np.float32)
        g = np.ascontiguousarray(np.atleast_1d(gaps), np.int32)
        rid = C.c_int64()
        _check(lib().wm_server_submit(self.h, geom.h, len(f), _ptr(f, C.c_float), len(g), _ptr(g, C.c_int),
                                      float(ampl), float(phas), C.byref(rid)))
        self.meta[rid.value] = (len(f), len(g), geom.npls)
        return rid.value

    def step(self, wait: bool = False) -> int:
        """Retires finished batches and starts new ones; with wait, blocks until a request
        completes, one is submitted, or nothing is pending.  Returns the requests completed."""
        d = C.c_int()
        _check(lib().wm_server_step(self.h, int(wait), C.byref(d)))
        return d.value

    def overlap(self):
        """(batches, seconds with a lane busy, seconds with two or more lanes busy)."""
        b, busy, over = C.c_int64(), C.c_double(), C.c_double()
        _check(lib().wm_server_overlap(self.h, C.byref(b), C.byref(busy), C.byref(over)))
        return b.value, busy.value, over.value

    def done(self, cap: int = 4096):
        """Ids of completed requests whose results have not been read."""
        ids = (C.c_int64 * cap)()
        k = C.c_int()
        _check(lib().wm_server_done(self.h, ids, cap, C.byref(k)))
        return [ids[i] for i in range(k.value)]

    def status(self):
        q, r, c = C.c_int(), C.c_int(), C.c_int()
        used = C.c_size_t()
        _check(lib().wm_server_status(self.h, C.byref(q), C.byref(r), C.byref(c), C.byref(used)))
        return q.value, r.value, c.value, used.value

    def result(self, rid: int):
        nf, nr, n = self.meta[rid]
        I = np.zeros((nf, nr, n), np.complex128)
        zin = np.zeros((nf, nr), np.complex128)
        st = (FreqStats * nf)()
        lat = C.c_double()
        _check(lib().wm_server_result(self.h, rid, I.ctypes.data, zin.ctypes.data, st, C.byref(lat)))
        del self.meta[rid]
        return zin, I, [_stats(s) for s in st], lat.value

    def close(self):
        if self.h:
            lib().wm_server_destroy(self.h)
            self.h = None
