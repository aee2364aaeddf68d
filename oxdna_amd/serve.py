"""Served form: a request loop over resident slot pools on one GPU.

Jobs (one oxDNA2 system each, with its own temperature, seed and step
count) arrive over time.  The server keeps two kinds of slot pool resident
in HBM (include/oxh.h, oxh_pool_*):

  - a workgroup pool: slots of 128 nucleotides, one workgroup per job and
    many steps per launch;
  - grid pools: slots of a few capacity classes up to the largest job,
    one kernel per phase for all of a pool's jobs together, cell lists.

Admission is by HBM footprint: the pools are sized at start-up from a byte
budget (oxh_batch_stats reports what each pool holds), and a job is
admitted when a free slot of large enough capacity exists.  Each pool is
driven by its own host thread, so the pools' kernels overlap on the device;
every `quantum` steps a pool's thread releases finished jobs and admits
waiting ones into the freed slots while the other slots keep their state
(continuous batching).  A job's result does not depend on what else runs:
see oxh_pool_create.

The reference has no counterpart (one system per process,
src/Utilities/ConfigInfo.h:47); its input keys per job are honoured through
backend.make_desc.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .backend import HIPPool
from .systems import System

# jobs up to SMALL_CAP nucleotides run in the workgroup pool (one workgroup
# each, at least 4 lanes per nucleotide); a workgroup engine step costs ~n
# on one CU, so larger jobs would hold a CU for long: they go to grid pools
SMALL_CAP = 128


@dataclass
class Job:
    job_id: int
    system: System
    steps: int
    arrival: float = 0.0                 # seconds after the start of the run
    opts: dict = field(default_factory=dict)
    admitted: float | None = None
    finished: float | None = None
    error: int = 0
    result: dict | None = None

    @property
    def work(self) -> float:
        return float(self.system.N) * self.steps


def footprint_bytes(n_nucleotides: int, max_neigh: int = 64) -> int:
    """HBM held per nucleotide of slot capacity: FP64 state (pos, vel, L,
    quaternion, list position: 5 x 32 B), its chunk backup (5 x 32 B + 8),
    FP32 mirror, backbone sites, forces and torques (4 x 16 B), cells,
    bonds and info (~40 B), the neighbour row (4 B x max_neigh)."""
    return n_nucleotides * (160 + 168 + 64 + 40 + 4 * max_neigh)


def plan_pools(budget_bytes: int, largest: int, small_slots: int = 256):
    """Slot capacities for one workgroup pool and the grid pools, within
    budget_bytes: the workgroup pool takes small_slots x SMALL_CAP nt; the
    rest goes to grid slots in capacity classes x4 apart (512, 2048, ...)
    up to `largest`, about the same bytes per class."""
    small = [SMALL_CAP] * small_slots
    left = budget_bytes - footprint_bytes(SMALL_CAP * small_slots)
    classes = []
    c = 512
    while True:
        classes.append(min(c, largest))
        if c >= largest:
            break
        c *= 4
    per_class = max(left // max(1, len(classes)), 0)
    grid = []
    for c in classes:
        # every step of a grid pool launches over all its slots' capacity,
        # so a class holds at most ~64k nucleotides of slots
        n = max(1, int(per_class // footprint_bytes(c, 128)))
        grid.append([c] * max(1, min(n, 65536 // c)))
    return small, grid


class Server:
    """Resident pools and the request loop."""

    def __init__(self, budget_bytes: int = 8 << 30, largest: int = 20000, small_slots: int = 256, device: int = 0,
                 quantum_small: int = 1024, quantum_grid: int = 256):
        small, grid = plan_pools(budget_bytes, largest, small_slots)
        self.pools = [HIPPool(small, device=device)]
        self.quantum = [quantum_small]
        for caps in grid:
            self.pools.append(HIPPool(caps, device=device))
            self.quantum.append(quantum_grid)
        self.capacity = [max(p.capacity) for p in self.pools]
        self._lock = threading.Lock()

    def close(self):
        for p in self.pools:
            p.close()

    def hbm_bytes(self) -> int:
        return sum(p.stats()["device_bytes"] for p in self.pools)

    def _pool_for(self, n: int) -> int:
        for k, c in enumerate(self.capacity):
            if n <= c:
                return k
        raise ValueError(f"job of {n} nucleotides exceeds every slot")

    def run(self, jobs: list[Job], keep_results: bool = False) -> dict:
        """Serves `jobs` (arrival times in seconds from now); returns when
        all are done.  Each pool's thread admits, runs and releases."""
        queues = [[] for _ in self.pools]
        for j in sorted(jobs, key=lambda j: j.arrival):
            queues[self._pool_for(j.system.N)].append(j)
        t0 = time.perf_counter()
        errors = []

        def worker(k: int):
            pool, q, quantum = self.pools[k], queues[k], self.quantum[k]
            slots: list[Job | None] = [None] * pool.nsys
            caps = pool.capacity
            try:
                while True:
                    now = time.perf_counter() - t0
                    # admit arrived jobs into free slots (smallest fitting slot)
                    i = 0
                    while i < len(q) and q[i].arrival <= now:
                        j = q[i]
                        free = [s for s in range(pool.nsys) if slots[s] is None and caps[s] >= j.system.N]
                        if not free:
                            i += 1
                            continue
                        s = min(free, key=lambda s: caps[s])
                        pool.admit(s, j.system, j.steps, **j.opts)
                        j.admitted = time.perf_counter() - t0
                        slots[s] = j
                        q.pop(i)
                    if all(x is None for x in slots):
                        if not q:
                            return
                        wait = q[0].arrival - (time.perf_counter() - t0)
                        if wait > 0:
                            time.sleep(min(wait, 0.05))
                        continue
                    pool.run(quantum)
                    st = pool.status()
                    for s, j in enumerate(slots):
                        if j is None:
                            continue
                        if st["done"][s] >= st["target"][s]:
                            j.error = int(st["error"][s])
                            if keep_results:
                                j.result = pool.state(s)
                            pool.release(s)
                            j.finished = time.perf_counter() - t0
                            slots[s] = None
            except Exception as e:  # surfaced to the caller below
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(self.pools))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        t1 = max(j.finished for j in jobs)
        lat = np.array([j.finished - j.arrival for j in jobs])
        work = sum(j.work for j in jobs)
        return dict(makespan=t1, nt_steps_per_s=work / t1, work=work, p50=float(np.percentile(lat, 50)),
                    p99=float(np.percentile(lat, 99)), mean_latency=float(lat.mean()), n_jobs=len(jobs))


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times of n jobs of a Poisson stream of `rate` jobs/s."""
    if math.isinf(rate):
        return np.zeros(n)
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
