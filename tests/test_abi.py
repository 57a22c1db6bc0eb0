"""The C-ABI library loads and exports every function include/wiremom.h declares,
and the geometry front end matches the oracle bit for bit.  CPU only: no call
here touches a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "wiremom.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text,
                       flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def test_header_declares_the_drop_in_and_native_api():
    names = declared_functions()
    for n in ["zmatrix", "zmatrix_gpu", "CalculateByFreq", "LU_cuBLAS", "geoinp", "config", "configwire",
              "wm_init", "wm_zfill", "wm_getrf_batched", "wm_getrs_batched", "wm_solve_sweep", "wm_server_submit"]:
        assert n in names


def test_library_exports_every_declared_symbol(wm):
    L = wm.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(wm.EXPORTS) <= set(declared_functions())
    assert b"gfx950" in L.wm_version()


def test_library_has_no_cpu_fallback(wm):
    """The product library links no oracle code."""
    L = wm.lib()
    for sym in ["wmo_zmatrix", "wmo_zgetrf", "wmo_solve_freq"]:
        assert not hasattr(L, sym)


def _cmp_topology(t, to):
    ns = t.nsmax
    for a in ["x", "y", "z", "rad", "ipc"]:
        assert np.array_equal(getattr(t, a)[:ns], getattr(to, a)[:ns]), a
    for a in ["ise", "isp", "ips", "iwe"]:
        assert np.array_equal(getattr(t, a)[:2 * ns], getattr(to, a)[:2 * ns]), a
    assert (t.npnt, t.nseg, t.npls) == (to.npnt, to.nseg, to.npls)


@pytest.mark.parametrize("which", ["dipole768", "grid", "lightning", "cage_small", "tjunction",
                                   "monopole_pec", "lightning_pec"])
def test_front_end_matches_oracle(wm, geometries, oracle, geo_dir, which):
    """configwire in libwiremom (product) and in the oracle agree bit for bit:
    basis order, node ids and float32 coordinates, and configwire.txt text;
    the *_pec cases segment for the PEC ground (ienv 2, virtual ground wires)."""
    G = geometries
    p = os.path.join(geo_dir, f"fe_{which}.geo")
    ienv = 2 if which.endswith("_pec") else 1
    if which == "monopole_pec":
        G.monopole(p, 40)
    elif which == "lightning_pec":
        G.lightning(p, 1024, depth=0.0)
    elif which == "dipole768":
        G.dipole(p, 768)
    elif which == "grid":
        G.planar_grid(p, 1536, seed=11)
    elif which == "lightning":
        G.lightning(p, 2048)
    elif which == "cage_small":
        G.cage(p, 1200, seed=3, k=4)
    else:
        G.write_geo(p, [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, -1, 1)],
                    [(1, 2, 0.001, 10), (2, 3, 0.001, 10), (4, 2, 0.001, 10)])
    d = os.path.join(geo_dir, f"fe_{which}")
    os.makedirs(d, exist_ok=True)
    t = wm.configwire(p, ienv=ienv, cwd=d)
    to = oracle.load(p, ienv=ienv, wirefile=os.path.join(d, "oracle_cw.txt"))
    _cmp_topology(t, to)
    assert open(os.path.join(d, "configwire.txt"), "rb").read() == open(os.path.join(d, "oracle_cw.txt"), "rb").read()
    if ienv == 2:  # the virtual ground wires add bases
        assert t.npls > oracle.load(p, ienv=1).npls
