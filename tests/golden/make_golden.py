"""Writes the golden fixtures in tests/golden/ from the reference's own data files.

Run in a container where the reference checkout is mounted at /root/reference:

    python tests/golden/make_golden.py

It only reads text data the reference ships (run logs and configwire.txt
outputs) and stores numbers, never source text:

* zin_kat.json -- every "Process - r Freq = f Re Im" line of the
  results_P_F_N.txt logs: the 25 m straight dipole of dipol.geo (nodes
  (0,0,-12.5)/(0,0,12.5), radius 0.5 mm) cut into N segments, delta gap at
  basis 100, free space, Zin printed to 4 significant digits.  All four
  variants of the reference agree on every (N, f) pair.
* configwire_kat.json -- for each shipped configwire.txt: segment count,
  byte length and sha256 of the file, plus its first and last lines.
"""
import glob
import hashlib
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def zin_kat():
    rows = {}
    pat = re.compile(r"Freq = ([-\d.e+]+)\s+([-\d.e+]+)\s+([-\d.e+]+)\s*$")
    for f in sorted(glob.glob(os.path.join(REF, "*", "results_*.txt"))):
        m = re.match(r"results_(\d+)_(\d+)_(\d+)\.txt", os.path.basename(f))
        if not m:
            continue
        nseg = int(m.group(3))
        variant = os.path.basename(os.path.dirname(f))
        with open(f, encoding="latin-1") as fh:
            for line in fh:
                mm = pat.search(line)
                if mm:
                    key = (nseg, mm.group(1))
                    val = (mm.group(2), mm.group(3))
                    rows.setdefault(key, {}).setdefault(val, set()).add(variant)
    out = []
    for (nseg, f), vals in sorted(rows.items(), key=lambda kv: (kv[0][0], float(kv[0][1]))):
        # a line garbled by interleaved MPI output shows up as a minority value
        val = max(vals.items(), key=lambda kv: len(kv[1]))[0]
        out.append({"nseg": nseg, "freq_mhz": f, "re": val[0], "im": val[1],
                    "variants": sorted(vals[val])})
    return out


def configwire_kat():
    out = []
    for d in sorted(glob.glob(os.path.join(REF, "wersja*", "configwire.txt"))):
        data = open(d, "rb").read()
        lines = data.decode("latin-1").split("\n")
        m = re.search(r"functions: (\d+)\t(\d+)", lines[1])
        out.append({"variant": os.path.basename(os.path.dirname(d)), "nw": int(m.group(1)),
                    "npls": int(m.group(2)), "nseg": int(m.group(2)) + 1, "bytes": len(data),
                    "sha256": hashlib.sha256(data).hexdigest(), "head": lines[:6], "tail": lines[-4:]})
    return out


def main():
    kat = zin_kat()
    with open(os.path.join(HERE, "zin_kat.json"), "w") as fh:
        geometry = {"nodes": [[0, 0, -12.5], [0, 0, 12.5]], "radius": 0.0005, "ngap": 100, "ienv": 1}
        fh.write('{"geometry": %s,\n"cases": [\n' % json.dumps(geometry))
        fh.write(",\n".join(json.dumps(c) for c in kat))
        fh.write("\n]}\n")
    with open(os.path.join(HERE, "configwire_kat.json"), "w") as fh:
        json.dump(configwire_kat(), fh, indent=1)
    print(f"{len(kat)} Zin cases")


if __name__ == "__main__":
    main()
