"""ctypes front of the CPU oracle (oracle/wm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py`` for its ``cpu_baseline`` leg.  The product
(``libwiremom.so`` and the package around it) never imports this module.

Each function names the reference routine it restates; the C file carries the
file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# Capacities of the reference driver ("wersja referencyjna/Wire_MoM.cu":43-44).
NNMAX, NTMAX, NWMAX, NPWMAX = 3000, 5000, 10000, 20480
NSMAX = NWMAX + NPWMAX

_lib = None


def build() -> None:
    """Compile oracle/liboracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(t):
    return C.POINTER(t)


def _declare(L):
    fp, ip, dp = _p(C.c_float), _p(C.c_int), _p(C.c_double)
    L.wmo_geoinp.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, ip, fp, fp, fp, ip, ip, fp, ip, ip]
    L.wmo_read_freq.argtypes = [C.c_char_p, C.c_int, fp]
    L.wmo_configwire_path.argtypes = [C.c_int, fp, fp, fp, C.c_int, ip, fp, ip, C.c_int, C.c_int, C.c_int,
                                      fp, fp, fp, fp, ip, ip, ip, ip, ip, ip, ip, ip, C.c_char_p]
    L.wmo_delick.argtypes = [C.c_double]
    L.wmo_delick.restype = C.c_double
    L.wmo_ckmn.argtypes = [C.c_double, C.c_int, C.c_double] + [C.c_double] * 11 + [dp, dp]
    L.wmo_cknn.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
    L.wmo_self_integral.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
    L.wmo_zmatrix_cols.argtypes = [dp, C.c_int, fp, fp, fp, ip, ip, ip, fp, C.c_int, C.c_double, C.c_int,
                                   C.c_int, C.c_int]
    L.wmo_zmatrix_env_cols.argtypes = [dp, C.c_int, fp, fp, fp, ip, ip, ip, fp, C.c_int, C.c_double, C.c_int,
                                       C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]
    L.wmo_wavenumber.argtypes = [C.c_float]
    L.wmo_wavenumber.restype = C.c_double
    L.wmo_zgetrf.argtypes = [C.c_int, dp, C.c_int, ip]
    L.wmo_zgetrs.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, dp, C.c_int]
    L.wmo_cdiv.argtypes = [C.c_double] * 4 + [dp, dp]
    L.wmo_num_threads.restype = C.c_int
    L.wmo_set_threads.argtypes = [C.c_int]
    L.wmo_tablice.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, fp, fp, fp, C.c_int, ip, ip, ip]
    L.wmo_plane_wave_rhs.argtypes = [fp, fp, fp, ip, ip, ip, C.c_int, C.c_int] + [C.c_double] * 7 + [dp]
    L.wmo_plane_wave_rhs.restype = None


def _ptr(a, t):
    return a.ctypes.data_as(_p(t))


def set_threads(n: int) -> None:
    lib().wmo_set_threads(int(n))


def num_threads() -> int:
    return lib().wmo_num_threads()


@dataclass
class GeoInput:
    """Node and wire tables of a .geo file, 1-based as in geoinp."""
    nnod: int
    xnod: np.ndarray
    ynod: np.ndarray
    znod: np.ndarray
    nw: int
    iwn: np.ndarray
    wrad: np.ndarray
    nsw: np.ndarray


@dataclass
class Topology:
    """Arrays after configwire, 0-based positions, 1-based stored ids."""
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


def read_geo(path: str) -> GeoInput:
    """geoinp ("wersja referencyjna/Configuration.cpp":5-75)."""
    L = lib()
    xn = np.zeros(NNMAX + 2, np.float32)
    yn = np.zeros_like(xn)
    zn = np.zeros_like(xn)
    iwn = np.zeros(2 * NWMAX + 2, np.int32)
    wrad = np.zeros(NWMAX + 2, np.float32)
    nsw = np.zeros(NWMAX + 2, np.int32)
    nnod, nw, nt = C.c_int(), C.c_int(), C.c_int()
    rc = L.wmo_geoinp(path.encode(), NNMAX, NWMAX, NTMAX, C.byref(nnod), _ptr(xn, C.c_float),
                      _ptr(yn, C.c_float), _ptr(zn, C.c_float), C.byref(nw), _ptr(iwn, C.c_int),
                      _ptr(wrad, C.c_float), _ptr(nsw, C.c_int), C.byref(nt))
    if rc != 0:
        raise ValueError(f"geoinp({path}) failed with code {rc}")
    nw_ = nw.value
    # the reference stores the second node of wire i at iwn[i+nw]; repack to a
    # dense 1-based layout with that same stride
    return GeoInput(nnod.value, xn, yn, zn, nw_, iwn[: 2 * nw_ + 1].copy(), wrad, nsw)


def read_freq(path: str, maxn: int = 100000) -> np.ndarray:
    """freq.txt reader ("wersja referencyjna/Wire_MoM.cu":253-265): float32 MHz."""
    buf = np.zeros(maxn, np.float32)
    n = lib().wmo_read_freq(path.encode(), maxn, _ptr(buf, C.c_float))
    if n < 0:
        raise ValueError(f"cannot read {path}")
    return buf[:n].copy()


def configwire(g: GeoInput, ienv: int = 1, nsmax: int = NSMAX, nwmax: int = NWMAX,
               wirefile: str | None = None) -> Topology:
    """configwire ("wersja referencyjna/Configuration.cpp":112-520)."""
    L = lib()
    x = np.zeros(nsmax + 2, np.float32)
    y = np.zeros_like(x)
    z = np.zeros_like(x)
    rad = np.zeros_like(x)
    ise = np.zeros(2 * nsmax + 2, np.int32)
    isp = np.zeros_like(ise)
    ips = np.zeros_like(ise)
    iwe = np.zeros_like(ise)
    ipc = np.zeros(nsmax + 2, np.int32)
    npnt, nseg, npls = C.c_int(), C.c_int(), C.c_int()
    rc = L.wmo_configwire_path(g.nnod, _ptr(g.xnod, C.c_float), _ptr(g.ynod, C.c_float),
                               _ptr(g.znod, C.c_float), g.nw, _ptr(g.iwn, C.c_int),
                               _ptr(g.wrad, C.c_float), _ptr(g.nsw, C.c_int), ienv, nsmax, nwmax,
                               _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(z, C.c_float),
                               _ptr(rad, C.c_float), _ptr(ise, C.c_int), _ptr(isp, C.c_int),
                               _ptr(ips, C.c_int), _ptr(ipc, C.c_int), _ptr(iwe, C.c_int),
                               C.byref(npnt), C.byref(nseg), C.byref(npls),
                               (wirefile or "").encode())
    if rc != 0:
        raise ValueError(f"configwire failed with code {rc}")
    return Topology(nsmax, npnt.value, nseg.value, npls.value, x, y, z, rad, ise, isp, ips, ipc, iwe)


def tablice(t: Topology, path: str) -> None:
    """tablice ("wersja referencyjna/Configuration.cpp":524-556): geometry tables to `path`."""
    L = lib()
    rc = L.wmo_tablice(path.encode(), t.npnt, t.nseg, t.nsmax, _ptr(t.x, C.c_float), _ptr(t.y, C.c_float),
                       _ptr(t.z, C.c_float), t.npls, _ptr(t.ise, C.c_int), _ptr(t.ips, C.c_int),
                       _ptr(t.ipc, C.c_int))
    if rc != 0:
        raise OSError(f"cannot write {path}")


def load(path: str, **kw) -> Topology:
    return configwire(read_geo(path), **kw)


def wavenumber(freq_mhz: float) -> float:
    """k from a float32 MHz value ("wersja referencyjna/Wire_MoM.cu":338-340)."""
    return lib().wmo_wavenumber(C.c_float(np.float32(freq_mhz)))


def delick(bet: float) -> float:
    return lib().wmo_delick(float(bet))


def cknn(s, ifun, ka, dl, an) -> complex:
    re, im = C.c_double(), C.c_double()
    lib().wmo_cknn(s, ifun, ka, dl, an, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def ckmn(s, ifun, ka, rc, rn, dc, an, dl) -> complex:
    re, im = C.c_double(), C.c_double()
    lib().wmo_ckmn(s, ifun, ka, *rc, *rn, *dc, an, dl, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def self_integral(ifun, ka, dl, an) -> complex:
    re, im = C.c_double(), C.c_double()
    lib().wmo_self_integral(ifun, ka, dl, an, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


# Environment arguments of zmatrix for ienv = 1 (free space) and 2 (PEC
# ground z = 0): sgnx, sgny, sgnz, sgnenv ("wersja referencyjna/Wire_MoM.cu":107-122).
ENV_SIGNS = {1: (1.0, 1.0, 1.0, 1.0), 2: (1.0, 1.0, -1.0, -1.0)}


def zmatrix(t: Topology, ka: float, cols: tuple[int, int] | None = None, ienv: int = 1) -> np.ndarray:
    """zmatrix ("wersja referencyjna/Zmatrix.cpp":27-52); ienv 2 adds the
    PEC-ground image terms (cwire1_im / cwire2_im, Zmatrix.cpp:342-446, 559-663).

    Returns Z as a Fortran-ordered complex128 (npls, npls) array; with
    ``cols=(n0, n1)`` only those columns are filled (others are zero)."""
    n = t.npls
    Z = np.zeros((n, n), np.complex128, order="F")
    n0, n1 = cols if cols else (0, n)
    sx, sy, sz, se = ENV_SIGNS[ienv]
    lib().wmo_zmatrix_env_cols(Z.ctypes.data_as(_p(C.c_double)), n, _ptr(t.x, C.c_float), _ptr(t.y, C.c_float),
                               _ptr(t.z, C.c_float), _ptr(t.ipc, C.c_int), _ptr(t.ips, C.c_int),
                               _ptr(t.ise, C.c_int), _ptr(t.rad, C.c_float), n, float(ka), t.nsmax, ienv,
                               sx, sy, sz, se, n0, n1)
    return Z


def zgetrf(A: np.ndarray):
    """LAPACK zgetrf semantics; returns (LU, ipiv 1-based, info)."""
    LU = np.array(A, dtype=np.complex128, order="F", copy=True)
    n = LU.shape[0]
    ipiv = np.zeros(n, np.int32)
    info = lib().wmo_zgetrf(n, LU.ctypes.data_as(_p(C.c_double)), n, _ptr(ipiv, C.c_int))
    return LU, ipiv, info


def zgetrs(LU: np.ndarray, ipiv: np.ndarray, b: np.ndarray) -> np.ndarray:
    n = LU.shape[0]
    B = np.array(b, dtype=np.complex128, order="F", copy=True)
    if B.ndim == 1:
        B = B.reshape(n, 1, order="F")
    lib().wmo_zgetrs(n, B.shape[1], LU.ctypes.data_as(_p(C.c_double)), n, _ptr(ipiv, C.c_int),
                     B.ctypes.data_as(_p(C.c_double)), n)
    return B if np.ndim(b) > 1 else B[:, 0]


def cdiv(a: complex, b: complex) -> complex:
    re, im = C.c_double(), C.c_double()
    lib().wmo_cdiv(a.real, a.imag, b.real, b.imag, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def excitation(npls: int, gaps, ampl: float = 1.0, phas: float = 0.0) -> np.ndarray:
    """Delta-gap right-hand sides ("wersja referencyjna/Wire_MoM.cu":300-309):
    one column per 1-based gap index, value ampl*exp(j*pi*phas/180)."""
    arg = np.pi * float(np.float32(phas)) / 180.0
    a = float(np.float32(ampl))
    gaps = np.atleast_1d(gaps)
    B = np.zeros((npls, len(gaps)), np.complex128, order="F")
    for c, g in enumerate(gaps):
        B[int(g) - 1, c] = complex(a * np.cos(arg), a * np.sin(arg))
    return B


def load_impedance(freq_mhz: float, r_ohm: float, l_uh: float, c_uf: float) -> complex:
    """Series load r + j(2 pi f l - 1/(2 pi f c)), f in MHz, l in uH, c in uF (c = 0: none);
    the reference's hard-wired loads use 1 + j 2 pi f 4.5 ("wersja referencyjna/Wire_MoM.cu":349)."""
    f = float(np.float32(freq_mhz))
    x = 2.0 * np.pi * f * l_uh
    if c_uf != 0.0:
        x -= 1.0 / (2.0 * np.pi * f * c_uf)
    return complex(r_ohm, x)


def plane_wave_rhs(t: Topology, freq_mhz: float, theta_deg: float, phi_deg: float, e_theta, e_phi) -> np.ndarray:
    """Tested incident plane-wave field per basis (wmo_plane_wave_rhs)."""
    b = np.zeros(t.npls, np.complex128)
    et, ep = complex(e_theta), complex(e_phi)
    lib().wmo_plane_wave_rhs(_ptr(t.x, C.c_float), _ptr(t.y, C.c_float), _ptr(t.z, C.c_float), _ptr(t.ipc, C.c_int),
                             _ptr(t.ips, C.c_int), _ptr(t.ise, C.c_int), t.npls, t.nsmax, wavenumber(freq_mhz),
                             float(theta_deg), float(phi_deg), et.real, et.imag, ep.real, ep.imag,
                             b.ctypes.data_as(C.POINTER(C.c_double)))
    return b


def solve_freq_ex(t: Topology, freq_mhz: float, nrhs: int, sources, loads=(), ienv: int = 1, plane_waves=()):
    """One frequency with generators (basis, rhs, ampl, phas), plane waves (rhs, theta,
    phi, e_theta, e_phi) and series loads
    (basis, r_ohm, l_uh, c_uf) added to Z's diagonal before the LU, as the
    reference's load block does ("wersja referencyjna/Wire_MoM.cu":344-363).
    Loads on one basis are summed first.  Returns (zin[nsrc] = V/I at each
    generator, I[npls, nrhs])."""
    ka = wavenumber(freq_mhz)
    Z = zmatrix(t, ka, ienv=ienv)
    per = {}
    for b, r, l, c in loads:
        per[int(b)] = per.get(int(b), 0j) + load_impedance(freq_mhz, r, l, c)
    for b, zl in per.items():
        Z[b - 1, b - 1] += zl
    B = np.zeros((t.npls, nrhs), np.complex128, order="F")
    vals = []
    for b, r, a, p in sources:
        v = excitation(1, [1], a, p)[0, 0]
        B[int(b) - 1, int(r)] += v
        vals.append(v)
    for r, th, ph, et, ep in plane_waves:
        B[:, int(r)] += plane_wave_rhs(t, freq_mhz, th, ph, et, ep)
    LU, ipiv, info = zgetrf(Z)
    if info:
        raise np.linalg.LinAlgError(f"zgetrf info={info}")
    X = zgetrs(LU, ipiv, B)
    zin = np.array([cdiv(v, X[int(b) - 1, int(r)]) for v, (b, r, _, _) in zip(vals, sources)])
    return zin, X


def solve_freq(t: Topology, freq_mhz: float, gaps=(100,), ampl: float = 1.0, phas: float = 0.0,
               ienv: int = 1):
    """One frequency of the reference loop ("wersja referencyjna/Wire_MoM.cu":335-395).

    Returns (zin[nrhs], I[npls, nrhs], Z)."""
    ka = wavenumber(freq_mhz)
    Z = zmatrix(t, ka, ienv=ienv)
    LU, ipiv, info = zgetrf(Z)
    if info:
        raise np.linalg.LinAlgError(f"zgetrf info={info}")
    B = excitation(t.npls, gaps, ampl, phas)
    X = zgetrs(LU, ipiv, B)
    gaps = np.atleast_1d(gaps)
    zin = np.array([cdiv(B[int(g) - 1, c], X[int(g) - 1, c]) for c, g in enumerate(gaps)])
    return zin, X, Z


def trapezoid_weights(freq_mhz) -> np.ndarray:
    """Weights (Hz) of the trapezoid rule over [0, f_0, ..., f_K-1] at f_0..f_K-1 (the node at 0 aside)."""
    f = np.asarray(freq_mhz, np.float32).astype(np.float64) * 1e6
    d = np.diff(np.concatenate([[0.0], f])) / 2.0
    w = d.copy()
    w[:-1] += d[1:]
    return w


def transient(freq_mhz, H, S, t) -> np.ndarray:
    """r[m, j] = f_0 Re X[0, m] + 2 sum_k w_k Re(X[k, m] e^{j 2 pi f_k t_j}), X = H S: the trapezoid rule
    for the inverse transform of a real response with its DC value taken as Re X(f_0) (the reference
    leaves this step out, "wersja referencyjna/Wire_MoM.cu":232-248, 390)."""
    f = np.asarray(freq_mhz, np.float32).astype(np.float64) * 1e6
    X = (1.0 if S is None else np.asarray(S))[:, None] * np.asarray(H).reshape(len(f), -1)
    C = trapezoid_weights(freq_mhz)[:, None] * X
    E = np.exp(2j * np.pi * np.outer(f, np.asarray(t, np.float64)))
    return f[0] * X[0].real[:, None] + 2.0 * (C.T[:, :, None] * E[None, :, :]).real.sum(axis=1)


def spectrum_dexp(freq_mhz, i0, alpha, beta) -> np.ndarray:
    w = 2 * np.pi * np.asarray(freq_mhz, np.float32).astype(np.float64) * 1e6
    return i0 * (1.0 / (alpha + 1j * w) - 1.0 / (beta + 1j * w))
