"""Request streams for the continuous-batching server (wm_server_*).

The reference solves one frequency per MPI rank in a loop
("wersja gpu/Wire_MoM.cu":331-339); a serving deployment instead receives
requests -- a structure, its frequencies and feeds -- at arbitrary times.
`poisson_requests` draws such a stream and `replay` plays it against a
Server: a submitter thread submits each request at its arrival time while the
calling thread drives the device with Server.step(), so requests that arrive
while others are on the GPU join the next batch.
"""
from __future__ import annotations

import threading
import time

import numpy as np


def poisson_requests(count: int, rate: float, seed: int = 5, sizes=(512, 1024, 2048, 4096, 8192),
                     fmin: float = 0.01, fmax: float = 50.0, max_freqs: int = 32, max_rhs: int = 8):
    """`count` requests with exponential inter-arrival times (mean 1/rate s).
    Each has N drawn uniformly from `sizes`, 1..max_freqs frequencies uniform in
    [fmin, fmax] MHz (float32, like freq.txt), 1..max_rhs feeds and a structure seed."""
    rng = np.random.RandomState(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        nf = int(rng.randint(1, max_freqs + 1))
        out.append({"t": t, "n": int(rng.choice(sizes)), "seed": 1000 + i,
                    "freqs": rng.uniform(fmin, fmax, nf).astype(np.float32),
                    "nrhs": int(rng.randint(1, max_rhs + 1))})
    return out


def replay(server, reqs, geoms, feeds, idle_sleep: float = 5e-4, progress_every: int = 0):
    """Submits reqs[i] (structure geoms[i], feeds feeds[i]) at its arrival time
    and steps the server until every request is complete; with progress_every,
    a line on stderr after every that many completions.

    Returns {"latency": [s per request], "results": [(zin, I, stats) per request],
    "makespan": s from the first arrival slot to the last completion}."""
    rid = [None] * len(reqs)
    err = []
    t0 = time.perf_counter()

    def submitter():
        try:
            for i, r in enumerate(reqs):
                d = t0 + r["t"] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                rid[i] = server.submit(geoms[i], r["freqs"], feeds[i])
        except Exception as e:  # surfaced by the driving thread
            err.append(e)

    th = threading.Thread(target=submitter, daemon=True)
    th.start()
    got = {}
    while len(got) < len(reqs):
        if err:
            th.join()
            raise err[0]
        q, r, _, _ = server.status()
        # block in the server while work is in flight (it wakes on a completion
        # or a submission); poll while nothing has arrived yet
        server.step(wait=q + r > 0)
        for k in server.done():
            got[k] = server.result(k)
            if progress_every and len(got) % progress_every == 0:
                import sys
                print(f"replay: {len(got)}/{len(reqs)} requests done at {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        if len(got) < len(reqs):
            q, r, _, _ = server.status()
            if q + r == 0:
                time.sleep(idle_sleep)
    makespan = time.perf_counter() - t0
    th.join()
    res = [got[rid[i]] for i in range(len(reqs))]
    return {"latency": [x[3] for x in res], "results": [x[:3] for x in res], "makespan": makespan}
