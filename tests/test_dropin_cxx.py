"""The reference program's own call sites compile against the library's C++
drop-in header (include/wiremom_dropin.hpp: the reference's parameter types,
`int &` outputs and its own complex struct) and link against libwiremom.so.

CPU: compile, link and run the front end (geoinp, config) on a dipole; npls
and the basis geometry match the oracle.  GPU: the same program runs
CalculateByFreq, zmatrix_gpu and LU_cuBLAS for one frequency; the Zin line and
the separate fill + LU agree with the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "praca-magisterska_amd")


@pytest.fixture(scope="module")
def program(tmp_path_factory, wm):
    wm.lib()  # the library is built
    d = tmp_path_factory.mktemp("dropin")
    exe = str(d / "reference_call_sites")
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "dropin", "reference_call_sites.cpp"), "-L", PKG, "-lwiremom",
           f"-Wl,-rpath,{PKG}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_reference_call_sites_compile_link_and_run_front_end(program, geometries, oracle, tmp_path):
    geo = str(tmp_path / "dipol.geo")
    geometries.dipole(geo, 512, 1.0, 0.001)
    r = subprocess.run([program], capture_output=True, text=True, cwd=str(tmp_path),
                       env={**os.environ, "WM_GEO_PATH": geo}, timeout=60)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("nnod")][0].split()
    vals = dict(zip(line[0::2], line[1::2]))
    t = oracle.load(geo)
    assert int(vals["npls"]) == t.npls == 511
    assert int(vals["nnod"]) == 2 and int(vals["nw"]) == 1
    assert os.path.exists(str(tmp_path / "configwire.txt"))


@pytest.mark.gpu
def test_reference_call_sites_solve(program, geometries, oracle, tmp_path):
    geo = str(tmp_path / "dipol.geo")
    geometries.dipole(geo, 512, 1.0, 0.001)
    r = subprocess.run([program, "solve"], capture_output=True, text=True, cwd=str(tmp_path),
                       env={**os.environ, "WM_GEO_PATH": geo}, timeout=120)
    assert r.returncode == 0, r.stderr
    t = oracle.load(geo)
    zo, _, _ = oracle.solve_freq(t, 150.0, gaps=(256,))
    proc = [l for l in r.stdout.splitlines() if "Process - 0 Freq" in l][0].split()
    assert (proc[-2], proc[-1]) == ("%.4g" % zo[0].real, "%.4g" % zo[0].imag)
    zl = [l for l in r.stdout.splitlines() if l.startswith("zin")][0].split()
    z = complex(float(zl[1]), float(zl[2]))
    assert abs(z - zo[0]) <= 1e-10 * abs(zo[0])
