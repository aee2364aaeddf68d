"""Benchmark of the oxDNA2 MD hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain1]

Prints ONE JSON line.  A step is one full MD step of the workload: first
half kick + drift + rotation, on-device skin check and (when due) Verlet
rebuild, forces, second half kick, thermostat.

Workloads:
  plain1   512 random-sequence 24-bp duplexes (24,576 nt), cubic lattice,
           random orientations, density 1e-3 nt/unit^3, T = 25 C, 0.5 M,
           skin 0.05, dt 0.003, Bussi (newtonian_steps 10, bussi_tau 100)
           -- the headline (default)
  plain2   1,024 random-sequence 96-bp duplexes (196,608 nt), hexagonal
           lattice of axis spacing 2.6, T = 30 C, 0.2 M, Langevin
           (diff_coeff 2.5), skin 0.05, dt 0.003
  batched  256 hairpins of 16-60 nt in their own 30-unit boxes, 25-75 C,
           one launch for all systems
  served   a stream of jobs of 40-20,000 nt through resident slot pools
           (continuous batching), against the plain path one job at a time

value = nucleotide-steps/s of the whole job, inputs resident in HBM.
roofline: the force kernel (dominant), algorithmic bytes per launch /
average launch time (HIP events on the batch's stream).
cpu_baseline: the reference CPU binary (oracle/_ref/oxDNA_release, built
from the reference's own sources with its release flags) on a bounded
sample of the same system, single thread.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name: str, seed: int = 1):
    from oxdna_amd import systems
    if name == "plain1":
        s = systems.duplex_gas(seed, n_duplex=512, bp=24, density=1e-3, tc=25.0)
        opts = dict(thermostat="bussi", newtonian_steps=10, bussi_tau=100, dt=0.003, verlet_skin=0.05,
                    refresh_vel=True, seed=seed)
        ref_keys = dict(T="25C", salt_concentration="0.5", thermostat="bussi", newtonian_steps="10",
                        bussi_tau="100", box_type="cubic")
        desc = ("512 x 24-bp duplexes (24,576 nt), density 1e-3, 25 C, 0.5 M, Bussi, dt 0.003, skin 0.05")
        return s, opts, ref_keys, desc
    if name == "plain2":
        s = systems.duplex_bundle(seed, nx=32, ny=32, bp=96, spacing=2.6, lz=40.0, tc=30.0, salt=0.2)
        opts = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True, seed=seed)
        ref_keys = dict(T="30C", salt_concentration="0.2", thermostat="langevin", diff_coeff="2.5",
                        box_type="orthogonal")
        desc = ("1,024 x 96-bp duplexes (196,608 nt), hexagonal lattice 2.6, 30 C, 0.2 M, Langevin, dt 0.003")
        return s, opts, ref_keys, desc
    if name == "batched":
        hs = systems.hairpin_set(seed, n=256, box=30.0, t_lo=25.0, t_hi=75.0)
        opts = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True)
        per = [dict(seed=seed * 1000 + k) for k in range(len(hs))]
        desc = ("256 hairpins, 6-12 bp stems, 4-36 nt loops, 16-60 nt each, own 30-unit boxes, 25-75 C, 0.5 M, "
                "Langevin, dt 0.003")
        return hs, (opts, per), None, desc
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(s, ref_keys, steps_lo: int, steps_hi: int):
    """Wall time of the reference CPU binary for steps_hi - steps_lo steps
    (difference of two runs removes start-up), one thread."""
    from oxdna_amd import io
    binary = os.path.join(ROOT, "oracle", "_ref", "oxDNA_release")
    if not os.path.exists(binary):
        return None
    with tempfile.TemporaryDirectory() as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)

        def run(steps):
            keys = dict(backend="CPU", sim_type="MD", interaction_type="DNA2", steps=str(steps), seed="1", dt="0.003",
                        verlet_skin="0.05", refresh_vel="1", topology="t.top", conf_file="c.dat",
                        trajectory_file="traj.dat", energy_file="energy.dat", print_energy_every="100000000",
                        print_conf_interval="100000000", no_stdout_energy="1", restart_step_counter="1",
                        time_scale="linear", log_file="log.dat", lastconf_file="last.dat")
            keys.update(ref_keys)
            with open(os.path.join(d, "input"), "w") as f:
                for k, v in keys.items():
                    f.write(f"{k} = {v}\n")
            t0 = time.perf_counter()
            r = subprocess.run(["taskset", "-c", "0", binary, "input"], cwd=d, capture_output=True, text=True)
            if r.returncode != 0:
                r = subprocess.run([binary, "input"], cwd=d, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("reference CPU run failed")
            return time.perf_counter() - t0

        t_lo = run(steps_lo)
        t_hi = run(steps_hi)
    dt = t_hi - t_lo
    n = steps_hi - steps_lo
    return dict(value=s.N * n / dt, unit="nucleotide-steps/s", cores=1, kind="reference",
                sample=f"{n} MD steps of the same {s.N}-nt system (DNA2, release flags), "
                       f"{dt:.1f} s, ms/step {1e3 * dt / n:.2f}")


def served_main(args):
    """Served stream: jobs of log-uniform size 40-20,000 nt (single
    strands, duplexes, hairpins, bundles), T 20-60 C, Langevin, with step
    counts log-uniform over [1e4, 1e6] x step_scale.  First the plain path
    runs the jobs one at a time (grid engine, one batch per job, as the
    reference runs one system per process); then the server takes the same
    jobs as a Poisson stream at rate_mult x the plain path's job rate, and
    once more with every job arriving at t = 0 (capacity)."""
    from oxdna_amd import backend, serve, systems
    rng_seed = args.seed
    jobs = []
    for k in range(args.jobs):
        s, steps = systems.served_job(k, seed=rng_seed)
        steps = max(100, int(round(steps * args.step_scale)))
        opts = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True,
                    seed=(rng_seed * 1000003 + k) & 0xFFFFFFFF)
        jobs.append(serve.Job(k, s, steps, opts=opts))
    work = sum(j.work for j in jobs)
    sizes = np.array([j.system.N for j in jobs])
    log(f"served: {len(jobs)} jobs, N {sizes.min()}-{sizes.max()} (median {int(np.median(sizes))}), "
        f"{work:.3e} nt-steps")

    # plain path, one job at a time
    os.environ["OXH_PERSIST"] = "0"
    t_setup = 0.0
    t_run = 0.0
    t0 = time.perf_counter()
    for j in jobs:
        ta = time.perf_counter()
        b = backend.HIPBatch(j.system, **j.opts)
        tb = time.perf_counter()
        b.run(j.steps)
        b.sync()
        tc = time.perf_counter()
        b.close()
        t_setup += tb - ta
        t_run += tc - tb
    t_plain = time.perf_counter() - t0
    del os.environ["OXH_PERSIST"]
    plain = work / t_plain
    log(f"plain one at a time: {t_plain:.2f} s ({t_setup:.2f} s upload/setup), {plain:.3e} nt-steps/s; "
        f"stepping only {work / t_run:.3e}")

    srv = serve.Server(budget_bytes=args.budget_gb << 30)
    hbm = srv.hbm_bytes()
    rate = args.rate_mult * len(jobs) / t_plain
    arr = serve.poisson_arrivals(len(jobs), rate, seed=rng_seed)
    for j, a in zip(jobs, arr):
        j.arrival = float(a)
    res = srv.run(jobs)
    log(f"served at {rate:.2f} jobs/s ({args.rate_mult}x plain): makespan {res['makespan']:.2f} s, "
        f"{res['nt_steps_per_s']:.3e} nt-steps/s, p50 {res['p50']:.3f} s p99 {res['p99']:.3f} s")
    for j in sorted(jobs, key=lambda j: -j.finished)[:5]:
        log(f"  job {j.job_id}: N {j.system.N} steps {j.steps} arrived {j.arrival:.2f} s admitted {j.admitted:.2f} s "
            f"finished {j.finished:.2f} s ({1e3 * (j.finished - j.admitted) / j.steps:.4f} ms/step in its pool)")
    for j in jobs:
        j.arrival = 0.0
        j.admitted = j.finished = None
    cap = srv.run(jobs)
    log(f"served, all at t=0: makespan {cap['makespan']:.2f} s, {cap['nt_steps_per_s']:.3e} nt-steps/s")
    srv.close()
    out = {
        "metric": "nucleotide-steps/s",
        "value": res["nt_steps_per_s"],
        "unit": "nucleotide-steps/s",
        "n_gpus": 1,
        "steps": int(sum(j.steps for j in jobs)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 state",
        "data": "synthetic",
        "config": {"workload": f"served stream: {len(jobs)} jobs, 40-20,000 nt log-uniform (single strands, duplexes, "
                               f"hairpins, bundles), 20-60 C, 0.5 M, Langevin, dt 0.003, steps log-uniform "
                               f"[1e4, 1e6] x {args.step_scale}, Poisson arrivals at {args.rate_mult}x the plain "
                               f"path's job rate",
                   "form": "served (slot pools, continuous batching)", "n_jobs": len(jobs), "work_nt_steps": work,
                   "arrival_rate_jobs_per_s": rate, "hbm_bytes_resident": hbm, "parallelism": "none"},
        "served": {"p50_latency_s": res["p50"], "p99_latency_s": res["p99"], "makespan_s": res["makespan"],
                   "capacity_nt_steps_per_s": cap["nt_steps_per_s"], "capacity_makespan_s": cap["makespan"]},
        "plain_one_at_a_time": {"value": plain, "seconds": t_plain, "setup_seconds": t_setup,
                                "stepping_only_value": work / t_run},
        "served_over_plain": res["nt_steps_per_s"] / plain,
        "capacity_over_plain": cap["nt_steps_per_s"] / plain,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--workload", default="plain1")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--sort-every", type=int, default=500, help="Hilbert re-sort interval in steps (0 = never)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "double"],
                    help="backend_precision: mixed (FP32 pair terms, FP64 state) or double (all FP64)")
    ap.add_argument("--jobs", type=int, default=64, help="served: number of jobs")
    ap.add_argument("--step-scale", type=float, default=0.01, help="served: factor on the jobs' step counts")
    ap.add_argument("--rate-mult", type=float, default=4.0, help="served: arrival rate / plain path's job rate")
    ap.add_argument("--budget-gb", type=int, default=8, help="served: HBM budget of the resident pools")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (systems are independent and a "
              "single system is not decomposed across GPUs in this build)", file=sys.stderr)
        sys.exit(2)

    from oxdna_amd import _lib, backend
    if _lib.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.workload == "served":
        return served_main(args)
    s, opts, ref_keys, desc = workload(args.workload, args.seed)
    if isinstance(s, list):
        common, per = opts
        systems_ = s
        b_args = dict(options=per, **common)
    else:
        systems_ = [s]
        b_args = dict(opts)
    b_args["backend_precision"] = args.precision
    n_total = sum(x.N for x in systems_)
    log(f"workload {args.workload}: {len(systems_)} system(s), N={n_total}")
    t_up = time.perf_counter()
    b = backend.HIPBatch(systems_, **b_args)
    b.set_sort_interval(args.sort_every)
    log(f"upload + first list/forces {time.perf_counter() - t_up:.2f} s")
    b.run(args.warmup)
    b.sync()
    reb0, _ = b.info()
    t0 = time.perf_counter()
    b.run(args.steps)
    b.sync()
    t1 = time.perf_counter()
    reb1, err = b.info()
    elapsed = t1 - t0
    value = n_total * args.steps / elapsed
    st = b.stats()
    force_ms = b.time_forces(50)
    E = b.energies().sum(axis=0)
    b.close()
    # achievable HBM bandwidth of a plain copy kernel on this device (and,
    # under a counter profiler, the dispatches that calibrate FETCH_SIZE /
    # WRITE_SIZE: 256 MiB fetched and written per launch)
    copy_gbps = backend.copy_bench(256 << 20, reps=3)
    nbar = st["mean_neigh"]
    # force kernel, algorithmic bytes per nucleotide: own state read
    # (pos 32, FP64 quaternion 32, backbone site 16, bonds 8, pinfo 8, list
    # lengths 8), F and T written (32), one 4-byte list entry per
    # neighbour; neighbour gathers are counted as cache-resident
    # (all-FP64 mode: F and T are written as FP64, 32 more bytes)
    b_alg = 136.0 + 4.0 * nbar + (32.0 if args.precision == "double" else 0.0)
    achieved = n_total * b_alg / (force_ms * 1e-3) / 1e9
    rebuild_every = args.steps / max(1, reb1 - reb0)
    log(f"{elapsed:.3f} s for {args.steps} steps; {1e3 * elapsed / args.steps:.4f} ms/step; force kernel "
        f"{force_ms * 1e3:.1f} us; mean neighbours {nbar:.1f}; rebuild every {rebuild_every:.1f} steps; "
        f"U/N {E[0] / n_total:.4f} K/N {(E[2] + E[3]) / n_total:.4f}")

    cpu = None
    if not args.no_cpu and ref_keys is not None:
        try:
            cpu = cpu_baseline(s, ref_keys, 20, 20 + args.cpu_steps)
        except Exception as e:  # the GPU figure stands on its own
            log(f"cpu baseline failed: {e}")

    out = {
        "metric": "nucleotide-steps/s",
        "value": value,
        "unit": "nucleotide-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 state" if args.precision == "mixed" else "f64 forces / f64 state",
        "data": "synthetic",
        "config": {"workload": desc, "n_nucleotides": n_total, "n_systems": len(systems_),
                   "backend_precision": args.precision,
                   "form": "plain (one system)" if len(systems_) == 1 else "batched (one launch for all systems)",
                   "parallelism": "none",
                   "hilbert_sort_every_steps": args.sort_every,
                   "mean_neighbours": round(nbar, 2), "steps_per_rebuild": round(rebuild_every, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "copy_kernel_gbps": copy_gbps,
                     "kernel": "k_forces" if args.precision == "mixed" else "k_forces_dp", "avg_launch_us": force_ms * 1e3, "bytes_per_nt": b_alg,
                     "lanes_per_nt": st["lanes"], "block": st["block"], "far_lanes_per_nt": st["far_lanes"],
                     "far_block": st["far_block"], "packed": st["pack"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
