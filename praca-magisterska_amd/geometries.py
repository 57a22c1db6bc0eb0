"""Synthetic wire structures written as .geo files (the reference's input format).

The reference reads ``dipol.geo`` ("wersja referencyjna/Configuration.cpp":5-75):
two comment lines, the node count and ``x y z`` per node, the wire count and
``node1 node2 radius nseg`` per wire, then a patch count.

The number of unknowns produced by configwire for a connected structure in
which every node ends at least one wire is

    N = sum(nseg) - nwires + (2 * nwires - nodes) = sum(nseg) + nwires - nodes

(each wire contributes nseg-1 interior basis functions, and each wire end
that meets an earlier wire's end adds one junction basis).  The generators
below pick per-wire segment counts so that N hits its target exactly.

Structures (sizes as used by the benchmarks and tests):
  dipole          straight wire along z (the reference's dipol.geo family)
  planar_grid     jittered planar lattice in z = 0, random segment lengths
  lightning       4 down-conductors, roof mesh at 1 m pitch, ground ring
                  (depth=0 puts the ring on the PEC ground plane z = 0)
  monopole        vertical wire standing on the ground plane
  cage            jittered cubic lattice of wires
"""
from __future__ import annotations

import math

import numpy as np


def write_geo(path: str, nodes, wires, title: str = "wire structure") -> None:
    """nodes: (V,3) float; wires: list of (n1, n2, radius, nseg) with 1-based node ids."""
    with open(path, "w") as f:
        f.write(f"{title}\n\n{len(nodes)}\n")
        for x, y, z in nodes:
            f.write(f"{float(np.float32(x))!r}\t{float(np.float32(y))!r}\t{float(np.float32(z))!r}\n")
        f.write(f"{len(wires)}\n")
        for n1, n2, a, ns in wires:
            f.write(f"{int(n1)}\t{int(n2)}\t{float(np.float32(a))!r}\t{int(ns)}\n")
        f.write("0\n")


def unknowns(nodes, wires) -> int:
    used = set()
    for n1, n2, _, _ in wires:
        used.add(n1)
        used.add(n2)
    return int(sum(w[3] for w in wires) + len(wires) - len(used))


def dipole(path: str, nseg: int, length: float = 25.0, radius: float = 0.0005) -> int:
    """Straight wire from (0,0,-L/2) to (0,0,L/2): N = nseg - 1."""
    nodes = [(0.0, 0.0, -length / 2), (0.0, 0.0, length / 2)]
    wires = [(1, 2, radius, nseg)]
    write_geo(path, nodes, wires, "Dipole")
    return nseg - 1


def monopole(path: str, nseg: int, height: float = 1.0, radius: float = 0.001) -> None:
    """Vertical wire from the ground plane (0,0,0) to (0,0,height); with the PEC
    ground (ienv 2) configwire adds the virtual ground wire at its base."""
    write_geo(path, [(0.0, 0.0, 0.0), (0.0, 0.0, height)], [(1, 2, radius, nseg)], "Monopole")


def _fit_nseg(lengths, target_sum, lo_len, hi_len, rng):
    """Integer segment counts with sum target_sum and length/nseg in [lo_len, hi_len] where possible."""
    lengths = np.asarray(lengths, float)
    seg = rng.uniform(lo_len, hi_len, size=len(lengths))
    ns = np.maximum(1, np.round(lengths / seg)).astype(int)
    scale = ns.sum() / target_sum
    ns = np.maximum(1, np.round(ns / scale)).astype(int)
    lo_n = np.maximum(1, np.ceil(lengths / hi_len)).astype(int)
    hi_n = np.maximum(lo_n, np.floor(lengths / lo_len)).astype(int)
    ns = np.clip(ns, lo_n, hi_n)
    order = rng.permutation(len(lengths))
    guard = 0
    while ns.sum() != target_sum and guard < 100 * len(lengths):
        d = 1 if ns.sum() < target_sum else -1
        moved = False
        for i in order:
            if ns.sum() == target_sum:
                break
            if (d > 0 and ns[i] < hi_n[i]) or (d < 0 and ns[i] > lo_n[i]):
                ns[i] += d
                moved = True
        if not moved:  # range exhausted: relax the bound
            i = order[guard % len(order)]
            ns[i] = max(1, ns[i] + d)
        guard += 1
    return ns


def planar_grid(path: str, n_target: int = 4096, seed: int = 2, size: float = 10.0, radius: float = 0.004,
                seg_lo: float = 0.025, seg_hi: float = 0.05, jitter: float = 0.1) -> int:
    """Jittered k x k lattice in z=0 spanning size x size m; returns N (== n_target)."""
    rng = np.random.RandomState(seed)
    mean_seg = 0.5 * (seg_lo + seg_hi)
    # lattice order so that the wire length at mean segment length gives about n_target
    k = 2
    while True:
        length = 2 * size * k
        if length / mean_seg * 0.92 + 2 * k * (k - 1) - k * k >= n_target or k > 60:
            break
        k += 1
    pitch = size / (k - 1)
    xy = np.array([(i * pitch, j * pitch) for j in range(k) for i in range(k)], float)
    xy += rng.uniform(-jitter * pitch, jitter * pitch, xy.shape)
    nodes = [(x, y, 0.0) for x, y in xy]
    edges = []
    for j in range(k):
        for i in range(k):
            v = j * k + i + 1
            if i + 1 < k:
                edges.append((v, v + 1))
            if j + 1 < k:
                edges.append((v, v + k))
    lens = [float(np.linalg.norm(xy[a - 1] - xy[b - 1])) for a, b in edges]
    target_sum = n_target - len(edges) + len(nodes)
    ns = _fit_nseg(lens, target_sum, seg_lo, seg_hi, rng)
    wires = [(a, b, radius, int(n)) for (a, b), n in zip(edges, ns)]
    write_geo(path, nodes, wires, f"Planar grid seed {seed}")
    n = unknowns(nodes, wires)
    assert n == n_target, (n, n_target)
    return n


def lightning(path: str, n_target: int = 8192, radius: float = 0.005, height: float = 20.0, depth: float = 0.5,
              size: float = 10.0, pitch: float = 1.0, ring: bool = True) -> int:
    """Lightning-protection structure: roof mesh at z=height, four corner down-conductors
    to z=-depth and (ring=True) a ring joining their bottoms.  Down-conductor 1 starts
    at (0,0,height).  With depth=0 and ring=False the down-conductors end on the PEC
    ground plane (grounded rods); a ring lying in the plane would be shorted by its
    image (zero columns of Z under ienv 2)."""
    k = int(round(size / pitch)) + 1
    nodes = [(i * pitch, j * pitch, height) for j in range(k) for i in range(k)]
    corners = [1, k, k * (k - 1) + 1, k * k]
    bottoms = []
    for c in corners:
        x, y, _ = nodes[c - 1]
        nodes.append((x, y, -depth))
        bottoms.append(len(nodes))
    edges = [(corners[0], bottoms[0])] + [(c, b) for c, b in zip(corners[1:], bottoms[1:])]
    for j in range(k):
        for i in range(k):
            v = j * k + i + 1
            if i + 1 < k:
                edges.append((v, v + 1))
            if j + 1 < k:
                edges.append((v, v + k))
    b0, b1, b2, b3 = bottoms  # ring: (0,0)-(10,0)-(10,10)-(0,10)
    if ring:
        edges += [(b0, b1), (b1, b3), (b3, b2), (b2, b0)]
    P = np.array(nodes, float)
    lens = np.array([np.linalg.norm(P[a - 1] - P[b - 1]) for a, b in edges])
    target_sum = n_target - len(edges) + len(nodes)
    delta = lens.sum() / target_sum
    ns = np.maximum(1, np.round(lens / delta)).astype(int)
    i = 0
    while ns.sum() != target_sum:  # deterministic correction, longest wires first
        order = np.argsort(-lens)
        j = order[i % len(order)]
        ns[j] += 1 if ns.sum() < target_sum else -1
        ns[j] = max(ns[j], 1)
        i += 1
    wires = [(a, b, radius, int(n)) for (a, b), n in zip(edges, ns)]
    write_geo(path, nodes, wires, "Lightning protection structure")
    n = unknowns(nodes, wires)
    assert n == n_target, (n, n_target)
    return n


def cage(path: str, n_target: int = 14336, seed: int = 4, k: int = 8, pitch: float = 1.0, radius: float = 0.002,
         jitter: float = 0.02) -> int:
    """k x k x k jittered cubic lattice of wires."""
    rng = np.random.RandomState(seed)
    idx = lambda i, j, l: (l * k + j) * k + i + 1
    P = np.array([(i * pitch, j * pitch, l * pitch) for l in range(k) for j in range(k) for i in range(k)], float)
    P += rng.uniform(-jitter * pitch, jitter * pitch, P.shape)
    edges = []
    for l in range(k):
        for j in range(k):
            for i in range(k):
                v = idx(i, j, l)
                if i + 1 < k:
                    edges.append((v, idx(i + 1, j, l)))
                if j + 1 < k:
                    edges.append((v, idx(i, j + 1, l)))
                if l + 1 < k:
                    edges.append((v, idx(i, j, l + 1)))
    lens = np.array([np.linalg.norm(P[a - 1] - P[b - 1]) for a, b in edges])
    target_sum = n_target - len(edges) + len(P)
    base = target_sum // len(edges)
    ns = np.full(len(edges), base, int)
    extra = rng.permutation(len(edges))[: target_sum - ns.sum()]
    ns[extra] += 1
    wires = [(a, b, radius, int(n)) for (a, b), n in zip(edges, ns)]
    write_geo(path, P.tolist(), wires, f"Cubic cage seed {seed}")
    n = unknowns(P.tolist(), wires)
    assert n == n_target, (n, n_target)
    return n


def nearest_basis(t, point) -> int:
    """1-based index of the basis function whose peak node is nearest to `point`."""
    n = t.npls
    pk = np.abs(t.ipc[:n]) - 1
    d = (t.x[pk] - point[0]) ** 2 + (t.y[pk] - point[1]) ** 2 + (t.z[pk] - point[2]) ** 2
    return int(np.argmin(d)) + 1


# Frequencies of the benchmark configurations (MHz, float32 like freq.txt).
def freqs_c2():
    return np.linspace(1.0, 100.0, 12).astype(np.float32)


def freqs_c3():
    return np.logspace(math.log10(0.01), math.log10(10.0), 64).astype(np.float32)


def freqs_c4():
    return np.linspace(0.1, 50.0, 16).astype(np.float32)


def write_freq(path: str, freqs) -> None:
    """freq.txt: the count, then one value per line in MHz."""
    with open(path, "w") as f:
        f.write(f"{len(freqs)}\n")
        for v in freqs:
            f.write(f"{float(np.float32(v)):.8e}\n")
