"""Served form: a request loop over resident slot pools on one GPU.

Jobs (one oxDNA2 system each, with its own temperature, seed and step
count) arrive over time.  The server keeps three kinds of slot pool
resident in HBM (include/oxh.h, oxh_pool_*), one per size class:

  - small jobs (up to SMALL_CAP nucleotides): one workgroup pool, one
    workgroup per job and many steps per launch;
  - mid-size jobs: one grid pool with slots of two capacities, one kernel
    per phase for all of its jobs together, cell lists;
  - large jobs: single-slot grid pools, one per stream, so that the few
    jobs that hold most of the work (and set the makespan) step as fast as
    they would alone while the other pools fill the rest of the device.

Admission is by HBM footprint: the pools are sized at start-up from a byte
budget (oxh_batch_stats reports what each pool holds).  Each pool is driven
by its own host thread, so the pools' kernels overlap on the device; every
`quantum` steps a pool's thread releases finished jobs and takes waiting
jobs of its class into the freed slots while the other slots keep their
state (continuous batching), longest waiting job first.  A job's result does not
depend on what else runs: see oxh_pool_create.

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

# jobs up to SMALL_CAP nucleotides run in the workgroup pool (one
# 512-thread workgroup each, at least 2 lanes per nucleotide: above 256
# nucleotides a workgroup has one lane per nucleotide, whose pair loop is a
# long chain of dependent gathers, and such jobs step faster in the grid
# pool); MID_CAP bounds the shared grid pool's slots; larger jobs get
# single-slot pools
SMALL_CAP = 256
MID_CAP = 8192


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


def plan_pools(budget_bytes: int, largest: int, small_slots: int = 256, mid_slots: tuple = (16, 4),
               large_pools: int = 2, small_cap: int = SMALL_CAP):
    """Pools as lists of slot capacities, one list per pool: the workgroup
    pool (small_slots x SMALL_CAP), the mid-size grid pool (mid_slots[0]
    slots of MID_CAP / 4 and mid_slots[1] of MID_CAP) and large_pools
    single-slot pools of `largest`; fewer mid-size slots if the budget
    is short.  Returns (pools, size class of each pool)."""
    small = [small_cap] * small_slots
    n1, n2 = mid_slots
    mid = [MID_CAP // 4] * n1 + [MID_CAP] * n2
    large = [[max(largest, MID_CAP)] for _ in range(large_pools)] if largest > MID_CAP else []
    fixed = footprint_bytes(small_cap * small_slots) + sum(footprint_bytes(p[0], 128) for p in large)
    while len(mid) > 1 and fixed + footprint_bytes(sum(mid), 128) > budget_bytes:
        mid.pop(0)
    pools = [small, mid] + large
    classes = [small_cap, MID_CAP] + [p[0] for p in large]
    return pools, classes


class Server:
    """Resident pools and the request loop."""

    def __init__(self, budget_bytes: int = 8 << 30, largest: int = 20000, small_slots: int = 256, device: int = 0,
                 quantum_small: int = 1024, quantum_grid: int = 256, quantum_large: int = 1024, small_cap: int = SMALL_CAP):
        pools, classes = plan_pools(budget_bytes, largest, small_slots, small_cap=small_cap)
        self.pools = [HIPPool(caps, device=device) for caps in pools]
        self.classes = classes
        self.quantum = [quantum_small, quantum_grid] + [quantum_large] * (len(pools) - 2)
        self.capacity = [max(p.capacity) for p in self.pools]

    def close(self):
        for p in self.pools:
            p.close()

    def hbm_bytes(self) -> int:
        return sum(p.stats()["device_bytes"] for p in self.pools)

    def _class_of(self, n: int) -> int:
        for c in sorted(set(self.classes)):
            if n <= c:
                return c
        raise ValueError(f"job of {n} nucleotides exceeds every slot")

    def run(self, jobs: list[Job], keep_results: bool = False) -> dict:
        """Serves `jobs` (arrival times in seconds from now); returns when
        all are done.  Jobs wait in one queue per size class, in arrival
        order; each pool's thread admits from its class's queue, runs and
        releases."""
        queues = {c: [] for c in set(self.classes)}
        for j in sorted(jobs, key=lambda j: j.arrival):
            queues[self._class_of(j.system.N)].append(j)
        lock = threading.Lock()
        t0 = time.perf_counter()
        errors = []

        def take(q, fits, now):
            """Of the jobs of q that have arrived, the one with the most work
            (nucleotides x steps) that fits a free slot, under lock: the
            longest jobs set the makespan, so they start first."""
            with lock:
                best = None
                for i, j in enumerate(q):
                    if j.arrival > now:
                        break
                    s = fits(j.system.N)
                    if s is not None and (best is None or j.work > q[best[0]].work):
                        best = (i, s)
                if best is None:
                    return None
                j = q.pop(best[0])
                return j, best[1]

        def worker(k: int):
            pool, q, quantum = self.pools[k], queues[self.classes[k]], self.quantum[k]
            slots: list[Job | None] = [None] * pool.nsys
            caps = pool.capacity

            def fits(n):
                free = [s for s in range(pool.nsys) if slots[s] is None and caps[s] >= n]
                return min(free, key=lambda s: caps[s]) if free else None

            try:
                while True:
                    # admit arrived jobs into free slots (smallest fitting slot)
                    while True:
                        got = take(q, fits, time.perf_counter() - t0)
                        if got is None:
                            break
                        j, s = got
                        pool.admit(s, j.system, j.steps, **j.opts)
                        j.admitted = time.perf_counter() - t0
                        slots[s] = j
                    if all(x is None for x in slots):
                        with lock:
                            if not q:
                                return
                            wait = q[0].arrival - (time.perf_counter() - t0)
                        time.sleep(min(max(wait, 0.0), 0.05) if wait > 0 else 0.001)
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
