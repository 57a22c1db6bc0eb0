"""The wire_mom driver: the reference program's I/O over the library.

"wersja referencyjna/Wire_MoM.cu":28-412 reads dipol.geo and freq.txt, writes
configwire.txt and prints " Process - r Freq = f<TAB>Re<TAB>Im" per frequency
with a sticky 4-significant-digit precision.  The CPU tests stop the driver at
the source-zone check (after the front end, before any GPU call); the GPU tests
run whole sweeps and compare the printed Zin with the oracle to the printed
precision.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "praca-magisterska_amd", "wire_mom")
LINE = re.compile(r"^ Process - (\d+) Freq = (\S+)\t(\S+)\t(\S+)$")


def _run(args, cwd, timeout=300):
    if not os.path.exists(DRIVER):
        pytest.fail("wire_mom is not built (python praca-magisterska_amd/build.py)")
    return subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def _lines(out):
    rows = []
    for l in out.splitlines():
        m = LINE.match(l)
        if m:
            rows.append((int(m.group(1)), float(m.group(2)), complex(float(m.group(3)), float(m.group(4)))))
    return rows


def test_usage_on_bad_flag(tmp_path):
    r = _run(["-x"], str(tmp_path))
    assert r.returncode == 2
    assert "usage" in r.stderr


@pytest.mark.parametrize("ienv", [1, 2])
def test_front_end_writes_reference_configwire(geometries, oracle, tmp_path, ienv):
    """configwire.txt from the driver equals the oracle's, byte for byte; a
    source zone outside 1..npls stops the run with the reference's message."""
    geo = str(tmp_path / "t.geo")
    if ienv == 1:
        geometries.planar_grid(geo, 1024, seed=7)
    else:
        geometries.monopole(geo, 60, 1.0, 0.001)
    r = _run(["-g", geo, "-e", str(ienv), "-s", "0"], str(tmp_path))
    assert r.returncode == 1
    assert "ERROR (wrong number of source zone)" in r.stdout
    assert " Nazwa pliku geo: " + geo in r.stdout
    want = str(tmp_path / "oracle_configwire.txt")
    oracle.configwire(oracle.read_geo(geo), ienv=ienv, wirefile=want)
    with open(tmp_path / "configwire.txt", "rb") as a, open(want, "rb") as b:
        assert a.read() == b.read()


def _write_freq(path, freqs, count=None):
    with open(path, "w") as f:
        f.write(f"{len(freqs) if count is None else count}\n")
        f.write("\n".join(str(x) for x in freqs) + "\n")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dipole_sweep_lines_match_oracle(geometries, oracle, tmp_path, prec):
    """C1's dipole (1 m, a = 1 mm, 512 segments) fed at its centre: every printed
    Zin equals the oracle's to the 4 significant digits printed."""
    geo = str(tmp_path / "dipol.geo")
    geometries.dipole(geo, 512, 1.0, 0.001)
    freqs = [150.0, 100.0, 300.0, 75.5]
    # the count may be smaller than the list, as in the reference's freq.txt
    _write_freq(tmp_path / "freq.txt", freqs + [999.0], count=len(freqs))
    r = _run(["-s", "256", "-P", prec, "-r", "3", "-t"], str(tmp_path))
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    rows = _lines(r.stdout)
    assert [f for _, f, _ in rows] == freqs
    assert all(rank == 3 for rank, _, _ in rows)
    t = oracle.load(geo)
    for _, f, z in rows:
        zo = oracle.solve_freq(t, f, gaps=(256,))[0][0]
        for got, want in ((z.real, zo.real), (z.imag, zo.imag)):
            assert float(f"{want:.4g}") == got, (f, z, zo)
    with open(tmp_path / "GPU time.dat") as fh:
        assert len(fh.read().strip().splitlines()) == 3


@pytest.mark.gpu
def test_monopole_over_ground_lines_match_oracle(geometries, oracle, tmp_path):
    """-e 2: a monopole on the PEC plane (image terms in the fill)."""
    geo = str(tmp_path / "mono.geo")
    geometries.monopole(geo, 200, 1.0, 0.001)
    freqs = [50.0, 75.0]
    _write_freq(tmp_path / "freq.txt", freqs)
    r = _run(["-g", geo, "-e", "2", "-s", "1"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = _lines(r.stdout)
    t = oracle.load(geo, ienv=2)
    assert len(rows) == len(freqs)
    for _, f, z in rows:
        zo = oracle.solve_freq(t, f, gaps=(1,), ienv=2)[0][0]
        assert abs(z - zo) <= 1e-3 * abs(zo), (f, z, zo)
        assert float(f"{zo.real:.4g}") == z.real and float(f"{zo.imag:.4g}") == z.imag


@pytest.mark.gpu
def test_loads_and_generators_files_match_oracle(geometries, oracle, tmp_path):
    """-L loads file and -S generators file: the printed Zin (V/I at the first
    generator) equals the oracle's to the 4 significant digits printed."""
    geo = str(tmp_path / "dipol.geo")
    geometries.dipole(geo, 512, 1.0, 0.001)
    freqs = [150.0, 300.0]
    _write_freq(tmp_path / "freq.txt", freqs)
    with open(tmp_path / "loads.txt", "w") as f:
        f.write("# basis R[ohm] L[uH] C[uF]\n100 50 0 0\n400 0 0.05 0\n400 10 0 1e-5\n")
    with open(tmp_path / "src.txt", "w") as f:
        f.write("256 1 0\n128 0.5 90  # second generator\n")
    r = _run(["-L", "loads.txt", "-S", "src.txt"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = _lines(r.stdout)
    assert [f for _, f, _ in rows] == freqs
    t = oracle.load(geo)
    loads = [(100, 50.0, 0.0, 0.0), (400, 0.0, 0.05, 0.0), (400, 10.0, 0.0, 1e-5)]
    for _, f, z in rows:
        zo = oracle.solve_freq_ex(t, f, 1, [(256, 0, 1.0, 0.0), (128, 0, 0.5, 90.0)], loads)[0][0]
        assert float(f"{zo.real:.4g}") == z.real and float(f"{zo.imag:.4g}") == z.imag, (f, z, zo)


def test_bad_loads_file_is_rejected(geometries, tmp_path):
    geo = str(tmp_path / "dipol.geo")
    geometries.dipole(geo, 64, 1.0, 0.001)
    _write_freq(tmp_path / "freq.txt", [100.0])
    with open(tmp_path / "loads.txt", "w") as f:
        f.write("10 50 0\n")
    r = _run(["-s", "32", "-L", "loads.txt"], str(tmp_path))
    assert r.returncode == 1
    assert "expected 4 numbers" in r.stderr


@pytest.mark.parametrize("ienv", [1, 2])
def test_tablice_matches_oracle(geometries, oracle, tmp_path, ienv):
    """-T writes tablice.txt ("wersja referencyjna/Configuration.cpp":524-556)
    byte for byte as the oracle's restatement (host only: the run stops at the
    source-zone check before any GPU call)."""
    geo = str(tmp_path / "t.geo")
    if ienv == 1:
        geometries.planar_grid(geo, 1024, seed=5)
    else:
        geometries.monopole(geo, 60, 1.0, 0.001)
    r = _run(["-g", geo, "-e", str(ienv), "-s", "0", "-T"], str(tmp_path))
    assert r.returncode == 1
    want = str(tmp_path / "oracle_tablice.txt")
    oracle.tablice(oracle.load(geo, ienv=ienv), want)
    with open(tmp_path / "tablice.txt", "rb") as a, open(want, "rb") as b:
        got, exp = a.read(), b.read()
    assert got == exp
    assert got.count(b" Point: ") > 100 if ienv == 1 else got.count(b" Point: ") > 10
