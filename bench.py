"""Benchmark of the oxDNA2 MD hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload headline]

Prints ONE JSON line.  A step is one full MD step of the workload: first
half kick + drift + rotation, on-device skin check and (when due) Verlet
rebuild, forces, second half kick, thermostat.

Workloads:
  headline (default)  the plain 196,608-nt system below in the mixed form
           (FP32 pair terms, FP64 state): `value`, `roofline` and
           `cpu_baseline`; the same system all-FP64 (`fp64`), the batched
           hairpins (`batched`), the batched oxRNA2 systems (`rna`, with
           the reference CPU binary on one of them) and the served stream
           against the plain path one job at a time (`served`) in the same
           line
  plain2   1,024 random-sequence 96-bp duplexes (196,608 nt), hexagonal
           lattice of axis spacing 2.6, box 83.2 x 72.05 x 40, T = 30 C,
           0.2 M, Langevin (diff_coeff 2.5), skin 0.05, dt 0.003
  plain1   512 random-sequence 24-bp duplexes (24,576 nt), cubic lattice,
           random orientations, density 1e-3 nt/unit^3, T = 25 C, 0.5 M,
           skin 0.05, dt 0.003, Bussi (newtonian_steps 10, bussi_tau 100)
  batched  256 hairpins of 16-60 nt in their own 30-unit boxes, 25-75 C,
           one launch for all systems
  served   a stream of jobs of 40-20,000 nt through resident slot pools
           (continuous batching), against the plain path one job at a time
  rna      oxRNA2: 24 systems of 1,024 nt, each 16 random-sequence 32-bp
           RNA duplexes in a box of 20, 37 C, 1 M, Langevin (diff_coeff
           2.5), skin 0.05, dt 0.003, one launch for all systems; the pair
           terms in FP64 (k_forces_rna)

value = nucleotide-steps/s of the whole job, inputs resident in HBM.
roofline: the force evaluation (the close-pair kernel k_forces and, when
the far pairs run on their own, k_forces_far; one launch each per step),
timed with HIP events on the batch's stream.  Algorithmic bytes per
nucleotide-step: 208 (FP64 position, velocity, angular momentum and
quaternion, read and written) + 64 (FP32 mirror, written and read) + 8
(bonds) + 2 n (one 4-byte index per unordered neighbour pair) + (32 + 4 n)
/ k (a list rebuild every k steps), n = mean neighbours per nucleotide;
`achieved` divides them by the force evaluation's time, `step_achieved` by
the whole step's.  `traffic` (HBM bytes per force evaluation) and
`valu_busy` come from the rocprofv3 counter passes in profiles/ for the
same workload and launch shape (calibrated by a plain copy kernel).
cpu_baseline: the reference CPU binary (oracle/_ref/oxDNA_release, built
from the reference's own sources with its release flags) on a bounded
sample of the same system, one thread per interaction variant (DNA2 and
DNA2_nomesh, on two cores at once).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import threading
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
        desc = ("plain 24,576-nt system: 512 x 24-bp random-sequence duplexes, density 1e-3, 25 C, 0.5 M, Bussi "
                "(newtonian_steps 10, bussi_tau 100), dt 0.003, skin 0.05")
        return s, opts, ref_keys, desc
    if name == "plain2":
        s = systems.duplex_bundle(seed, nx=32, ny=32, bp=96, spacing=2.6, lz=40.0, tc=30.0, salt=0.2)
        opts = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True, seed=seed)
        ref_keys = dict(T="30C", salt_concentration="0.2", thermostat="langevin", diff_coeff="2.5",
                        box_type="orthogonal")
        desc = ("plain 196,608-nt system: 1,024 x 96-bp random-sequence duplexes on a hexagonal lattice of axis "
                "spacing 2.6, box 83.2 x 72.05 x 40, 30 C, 0.2 M, Debye-Hueckel, Langevin (diff_coeff 2.5), dt 0.003, "
                "skin 0.05")
        return s, opts, ref_keys, desc
    if name == "rna":
        rs = systems.rna_duplex_set(seed, n_sys=24, per_side=4, bp=32, box=20.0, tc=37.0, salt=1.0)
        opts = dict(interaction_type="RNA2", thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05,
                    refresh_vel=True)
        per = [dict(seed=seed * 1000 + k) for k in range(len(rs))]
        desc = ("oxRNA2, 24 systems of 1,024 nt (24,576 nt): 16 random-sequence 32-bp RNA duplexes each on a 4 x 4 "
                "lattice in a cubic box of 20, 37 C, 1 M, Debye-Hueckel, Langevin (diff_coeff 2.5), dt 0.003, skin 0.05 "
                "(FP64 pair terms)")
        return rs, (opts, per), None, desc
    if name == "batched":
        hs = systems.hairpin_set(seed, n=256, box=30.0, t_lo=25.0, t_hi=75.0)
        opts = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True)
        per = [dict(seed=seed * 1000 + k) for k in range(len(hs))]
        desc = ("batched: 256 hairpins, 6-12 bp stems, 4-36 nt loops, 16-60 nt each, own 30-unit boxes, 25-75 C, "
                "0.5 M, Langevin, dt 0.003")
        return hs, (opts, per), None, desc
    raise SystemExit(f"unknown workload {name}")


def b_alg_per_nt_step(nbar: float, k: float) -> float:
    """Algorithmic HBM bytes per nucleotide-step of the whole MD step (see
    the module docstring)."""
    return 208.0 + 64.0 + 8.0 + 2.0 * nbar + (32.0 + 4.0 * nbar) / max(k, 1.0)


def profile_counters(workload_name: str, precision: str, st: dict):
    """Traffic and VALU figures of the committed counter profile for this
    workload, precision and launch shape (None if there is none)."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload_name}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    shape = c.get("shape", {})
    same = all(shape.get(k) == st.get(k) for k in ("lanes", "far_lanes", "pack"))
    return dict(c, path=os.path.relpath(path, ROOT), same_shape=same)


def run_plain(name: str, precision: str, steps: int, warmup: int, sort_every: int, seed: int):
    """Times exactly `steps` MD steps after `warmup` untimed ones; returns
    the figures of the run."""
    from oxdna_amd import backend
    s, opts, ref_keys, desc = workload(name, seed)
    if isinstance(s, list):
        common, per = opts
        systems_ = s
        b_args = dict(options=per, **common)
    else:
        systems_ = [s]
        b_args = dict(opts)
    b_args["backend_precision"] = precision
    n_total = sum(x.N for x in systems_)
    log(f"workload {name} ({precision}): {len(systems_)} system(s), N={n_total}")
    t_up = time.perf_counter()
    b = backend.HIPBatch(systems_, **b_args)
    b.set_sort_interval(sort_every)
    log(f"upload + launch-shape timing + first list/forces {time.perf_counter() - t_up:.2f} s")
    b.run(warmup)
    b.sync()
    reb0 = sum(b.info(k)[0] for k in range(len(systems_)))
    t0 = time.perf_counter()
    b.run(steps)
    b.sync()
    t1 = time.perf_counter()
    reb1 = sum(b.info(k)[0] for k in range(len(systems_)))
    elapsed = t1 - t0
    st = b.stats()
    force_ms = b.time_forces(50)
    E = b.energies().sum(axis=0)
    b.close()
    rebuilds = max(0, reb1 - reb0) / len(systems_)
    k = steps / rebuilds if rebuilds > 0 else float(steps)
    out = dict(name=name, desc=desc, system=s if not isinstance(s, list) else None, ref_keys=ref_keys,
               n_total=n_total, n_systems=len(systems_), precision=precision, elapsed=elapsed,
               ms_per_step=1e3 * elapsed / steps, value=n_total * steps / elapsed, force_ms=force_ms,
               nbar=st["mean_neigh"], k=k, stats=st, U_per_nt=E[0] / n_total, K_per_nt=(E[2] + E[3]) / n_total)
    log(f"{elapsed:.4f} s for {steps} steps; {out['ms_per_step']:.4f} ms/step; force evaluation {force_ms * 1e3:.1f} us; "
        f"mean neighbours {out['nbar']:.1f}; a rebuild every {k:.1f} steps; U/N {out['U_per_nt']:.4f} "
        f"K/N {out['K_per_nt']:.4f}; lanes {st['lanes']} far lanes {st['far_lanes']} packed {st['pack']}")
    return out


def roofline(r: dict, copy_gbps, counters):
    """roofline object of run r: the force evaluation against HBM peak."""
    st = r["stats"]
    b_alg = b_alg_per_nt_step(r["nbar"], r["k"])
    bytes_launch = b_alg * r["n_total"]
    achieved = bytes_launch / (r["force_ms"] * 1e-3) / 1e9
    step_achieved = bytes_launch / (r["ms_per_step"] * 1e-3) / 1e9
    if r["name"] == "rna":
        kernel = "k_forces_rna"
    elif r["precision"] == "double":
        kernel = "k_forces_dp" + (" + k_forces_far_dp" if st["far_lanes"] else "")
    else:
        kernel = "k_forces" + (" + k_forces_far" if st["far_lanes"] else "")
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": None, "kernel": kernel, "avg_launch_us": r["force_ms"] * 1e3,
           "bytes_per_nt_step": b_alg, "algorithmic_bytes_per_launch": bytes_launch,
           "mean_neighbours": r["nbar"], "steps_per_rebuild": r["k"],
           "step_achieved": step_achieved, "step_frac": step_achieved / HBM_PEAK_GBS,
           "copy_kernel_gbps": copy_gbps, "lanes_per_nt": st["lanes"], "block": st["block"],
           "far_lanes_per_nt": st["far_lanes"], "far_block": st["far_block"], "packed": st["pack"]}
    if counters is not None:
        out["traffic"] = counters["traffic_bytes_per_launch"] if counters["same_shape"] else None
        out["traffic_over_algorithmic"] = (counters["traffic_bytes_per_launch"] / bytes_launch
                                           if counters["same_shape"] else None)
        out["valu_busy"] = counters.get("valu_busy")
        out["counters_source"] = counters["path"] + ("" if counters["same_shape"] else " (other launch shape: not used)")
    return out


def rna_cpu_baseline(seed: int):
    """The reference CPU binary (RNA2, one thread) on one of the rna
    workload's 24 systems for 100 steps; value = the whole workload's
    nucleotide-steps/s at that rate (24 systems one after another)."""
    rs, opts, _, _ = workload("rna", seed)
    ref = dict(interaction_type="RNA2", T="37C", salt_concentration="1.0", thermostat="langevin", diff_coeff="2.5",
               box_type="cubic")
    c = cpu_baselines(rs[0], ref, 10, 110, variants=("RNA2",))
    if c is None:
        return None
    ms = c["variants"]["RNA2"]["ms_per_step"] * len(rs)
    n = sum(x.N for x in rs)
    return {"value": n / (ms * 1e-3), "unit": "nucleotide-steps/s", "cores": 1, "kind": "reference",
            "sample": (f"100 MD steps of one 1,024-nt system of the workload (difference of a 110- and a 10-step run), "
                       f"one thread, RNA2: {c['variants']['RNA2']['ms_per_step']:.2f} ms/step, x {len(rs)} systems"),
            "cpu_model": c["cpu_model"]}


def cpu_baselines(s, ref_keys, steps_lo: int, steps_hi: int, variants=("DNA2", "DNA2_nomesh")):
    """Wall time of the reference CPU binary for steps_hi - steps_lo steps
    (the difference of two runs removes start-up), one thread per variant,
    the variants on cores 0, 1, ... at the same time."""
    from oxdna_amd import io
    binary = os.path.join(ROOT, "oracle", "_ref", "oxDNA_release")
    if not os.path.exists(binary):
        return None
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    res = {}
    with tempfile.TemporaryDirectory() as d:
        io.write_topology(os.path.join(d, "t.top"), s)
        io.write_conf(os.path.join(d, "c.dat"), s)

        def run(variant, steps, core):
            w = os.path.join(d, f"{variant}_{steps}")
            os.makedirs(w)
            keys = dict(backend="CPU", sim_type="MD", interaction_type=variant, steps=str(steps), seed="1", dt="0.003",
                        verlet_skin="0.05", refresh_vel="1", topology="../t.top", conf_file="../c.dat",
                        trajectory_file="traj.dat", energy_file="energy.dat", print_energy_every="100000000",
                        print_conf_interval="100000000", no_stdout_energy="1", restart_step_counter="1",
                        time_scale="linear", log_file="log.dat", lastconf_file="last.dat")
            keys.update(ref_keys)
            with open(os.path.join(w, "input"), "w") as f:
                for k, v in keys.items():
                    f.write(f"{k} = {v}\n")
            t0 = time.perf_counter()
            r = subprocess.run(["taskset", "-c", str(core), binary, "input"], cwd=w, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"reference CPU run failed ({variant}): {r.stderr[-300:]}")
            return time.perf_counter() - t0

        def leg(variant, core):
            try:
                t_lo = run(variant, steps_lo, core)
                t_hi = run(variant, steps_hi, core)
                res[variant] = (t_lo, t_hi)
            except Exception as e:  # reported, the GPU figures stand on their own
                res[variant] = e

        threads = [threading.Thread(target=leg, args=(v, allowed[i % len(allowed)])) for i, v in enumerate(variants)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    n = steps_hi - steps_lo
    out = {}
    for v in variants:
        if isinstance(res.get(v), Exception) or v not in res:
            log(f"cpu baseline {v} failed: {res.get(v)}")
            continue
        t_lo, t_hi = res[v]
        dt = t_hi - t_lo
        out[v] = dict(value=s.N * n / dt, ms_per_step=1e3 * dt / n, seconds=dt)
    if not out:
        return None
    cpu_model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    main = out.get(variants[0]) or next(iter(out.values()))
    return {"value": main["value"], "unit": "nucleotide-steps/s", "cores": 1, "kind": "reference",
            "sample": (f"{n} MD steps of the same {s.N}-nt system per variant (difference of a {steps_hi}- and a "
                       f"{steps_lo}-step run), one thread each; value = {variants[0]} (the reference's default); "
                       + "; ".join(f"{v}: {o['ms_per_step']:.1f} ms/step" for v, o in out.items())),
            "variants": out, "cpu_model": cpu_model, "host_cores_visible": len(allowed),
            "binary": "oracle/_ref/oxDNA_release (reference CPU sources, -O3 -ffast-math -march=x86-64-v3)"}


def served_measure(jobs_n: int, step_scale: float, rate_mult: float, budget_gb: int, seed: int, small_cap: int = 256,
                   quantum_small: int = 1024, more_rates=(8.0,)) -> dict:
    """Served stream: jobs of log-uniform size 40-20,000 nt (single
    strands, duplexes, hairpins, bundles), T 20-60 C, Langevin, with step
    counts log-uniform over [1e4, 1e6] x step_scale.  First the plain path
    runs the jobs one at a time (grid engine, one batch per job, as the
    reference runs one system per process); then the server takes the same
    jobs as a Poisson stream at rate_mult x the plain path's job rate, and
    once more with every job arriving at t = 0 (capacity)."""
    from oxdna_amd import backend, serve, systems
    jobs = []
    for k in range(jobs_n):
        s, steps = systems.served_job(k, seed=seed)
        steps = max(100, int(round(steps * step_scale)))
        opts = dict(thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, refresh_vel=True,
                    seed=(seed * 1000003 + k) & 0xFFFFFFFF)
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

    srv = serve.Server(budget_bytes=budget_gb << 30, small_cap=small_cap, quantum_small=quantum_small)
    hbm = srv.hbm_bytes()
    rate = rate_mult * len(jobs) / t_plain
    arr = serve.poisson_arrivals(len(jobs), rate, seed=seed)
    for j, a in zip(jobs, arr):
        j.arrival = float(a)
    res = srv.run(jobs)
    log(f"served at {rate:.2f} jobs/s ({rate_mult}x plain): makespan {res['makespan']:.2f} s, "
        f"{res['nt_steps_per_s']:.3e} nt-steps/s, p50 {res['p50']:.3f} s p99 {res['p99']:.3f} s")
    for j in sorted(jobs, key=lambda j: -j.finished)[:5]:
        log(f"  job {j.job_id}: N {j.system.N} steps {j.steps} arrived {j.arrival:.2f} s admitted {j.admitted:.2f} s "
            f"finished {j.finished:.2f} s ({1e3 * (j.finished - j.admitted) / j.steps:.4f} ms/step in its pool)")
    others = {}
    for m in more_rates:
        r2 = m * len(jobs) / t_plain
        for j, a in zip(jobs, serve.poisson_arrivals(len(jobs), r2, seed=seed)):
            j.arrival = float(a)
            j.admitted = j.finished = None
        o = srv.run(jobs)
        others[f"{m:g}x"] = {"value": o["nt_steps_per_s"], "served_over_plain": o["nt_steps_per_s"] / plain,
                             "p50_latency_s": o["p50"], "p99_latency_s": o["p99"], "makespan_s": o["makespan"]}
        log(f"served at {r2:.2f} jobs/s ({m}x plain): makespan {o['makespan']:.2f} s, {o['nt_steps_per_s']:.3e} nt-steps/s "
            f"({o['nt_steps_per_s'] / plain:.2f}x), p50 {o['p50']:.3f} s p99 {o['p99']:.3f} s")
    for j in jobs:
        j.arrival = 0.0
        j.admitted = j.finished = None
    cap = srv.run(jobs)
    log(f"served, all at t=0: makespan {cap['makespan']:.2f} s, {cap['nt_steps_per_s']:.3e} nt-steps/s")
    srv.close()
    return {
        "workload": (f"served stream: {len(jobs)} jobs, 40-20,000 nt log-uniform (single strands, duplexes, hairpins, "
                     f"bundles), 20-60 C, 0.5 M, Langevin, dt 0.003, steps log-uniform [1e4, 1e6] x {step_scale}, "
                     f"Poisson arrivals at {rate_mult}x the plain path's job rate"),
        "value": res["nt_steps_per_s"], "unit": "nucleotide-steps/s", "n_jobs": len(jobs), "work_nt_steps": work,
        "step_scale": step_scale, "offered_load_vs_plain": rate_mult, "arrival_rate_jobs_per_s": rate,
        "p50_latency_s": res["p50"], "p99_latency_s": res["p99"], "makespan_s": res["makespan"],
        "capacity_nt_steps_per_s": cap["nt_steps_per_s"], "capacity_makespan_s": cap["makespan"],
        "plain_one_at_a_time": {"value": plain, "seconds": t_plain, "setup_seconds": t_setup,
                                "stepping_only_value": work / t_run},
        "served_over_plain": res["nt_steps_per_s"] / plain, "capacity_over_plain": cap["nt_steps_per_s"] / plain,
        "other_offered_loads": others,
        "hbm_bytes_resident": hbm, "workgroup_pool_slot_nt": small_cap, "workgroup_pool_quantum_steps": quantum_small,
    }


def line(r: dict, args, roof, cpu, extra=None) -> dict:
    out = {
        "metric": "nucleotide-steps/s",
        "value": r["value"],
        "unit": "nucleotide-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 state" if r["precision"] == "mixed" else "f64 forces / f64 state",
        "data": "synthetic",
        "config": {"workload": r["desc"], "n_nucleotides": r["n_total"], "n_systems": r["n_systems"],
                   "backend_precision": r["precision"],
                   "form": "plain (one system)" if r["n_systems"] == 1 else "batched (one launch for all systems)",
                   "parallelism": "none", "hilbert_sort_every_steps": args.sort_every,
                   "mean_neighbours": round(r["nbar"], 2), "steps_per_rebuild": round(r["k"], 2)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if extra:
        out.update(extra)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="headline", choices=["headline", "plain1", "plain2", "batched", "served", "rna"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="headline: skip the all-FP64 run")
    ap.add_argument("--no-served", action="store_true", help="headline: skip the served stream")
    ap.add_argument("--no-batched", action="store_true", help="headline: skip the batched hairpins")
    ap.add_argument("--no-rna", action="store_true", help="headline: skip the batched oxRNA2 systems")
    ap.add_argument("--cpu-steps", type=int, default=4,
                    help="CPU baseline: steps timed per variant (the 196,608-nt system takes seconds per step)")
    ap.add_argument("--sort-every", type=int, default=500, help="Hilbert re-sort interval in steps (0 = never)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "double"],
                    help="plain1/plain2/batched: backend_precision, mixed (FP32 pair terms, FP64 state) or double")
    ap.add_argument("--jobs", type=int, default=64, help="served: number of jobs")
    ap.add_argument("--step-scale", type=float, default=0.01, help="served: factor on the jobs' step counts")
    ap.add_argument("--rate-mult", type=float, default=4.0, help="served: arrival rate / plain path's job rate")
    ap.add_argument("--budget-gb", type=int, default=8, help="served: HBM budget of the resident pools")
    ap.add_argument("--small-cap", type=int, default=256, help="served: slot size of the workgroup pool (nucleotides)")
    ap.add_argument("--quantum-small", type=int, default=1024, help="served: steps between the workgroup pool's admissions")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (systems are independent and a "
              "single system is not decomposed across GPUs in this build)", file=sys.stderr)
        sys.exit(2)

    from oxdna_amd import _lib, backend
    if _lib.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.workload == "served":
        sv = served_measure(args.jobs, args.step_scale, args.rate_mult, args.budget_gb, args.seed, args.small_cap,
                            args.quantum_small)
        out = {"metric": "nucleotide-steps/s", "value": sv["value"], "unit": "nucleotide-steps/s", "n_gpus": 1,
               "steps": int(sv["work_nt_steps"]), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32 forces / f64 state", "data": "synthetic",
               "config": {"workload": sv["workload"], "form": "served (slot pools, continuous batching)",
                          "parallelism": "none"},
               "served": sv, "roofline": None, "cpu_baseline": None}
        print(json.dumps(out), flush=True)
        return

    name = "plain2" if args.workload == "headline" else args.workload
    precision = "mixed" if args.workload == "headline" else args.precision
    r = run_plain(name, precision, args.steps, args.warmup, args.sort_every, args.seed)
    # achievable HBM bandwidth of a plain copy kernel on this device (and,
    # under a counter profiler, the dispatches that calibrate FETCH_SIZE /
    # WRITE_SIZE: 256 MiB fetched and written per launch)
    copy_gbps = backend.copy_bench(256 << 20, reps=3)
    roof = roofline(r, copy_gbps, profile_counters(name, precision, r["stats"]))
    extra = {}
    if args.workload == "headline":
        if not args.no_fp64:
            r64 = run_plain(name, "double", args.steps, args.warmup, args.sort_every, args.seed)
            extra["fp64"] = {"value": r64["value"], "unit": "nucleotide-steps/s", "ms_per_step": r64["ms_per_step"],
                             "steps": args.steps, "warmup": args.warmup, "dtype": "f64 forces / f64 state",
                             "workload": r64["desc"] + " (backend_precision = double)",
                             "roofline": roofline(r64, copy_gbps, profile_counters(name, "double", r64["stats"]))}
        if not args.no_batched:
            # the workgroup engine runs up to 256 steps per launch: time whole launches
            nb = max(args.steps, 2560)
            rb = run_plain("batched", "mixed", nb, 256, args.sort_every, args.seed)
            extra["batched"] = {"value": rb["value"], "unit": "nucleotide-steps/s", "ms_per_step": rb["ms_per_step"],
                                "steps": nb, "warmup": 256, "dtype": "f32 forces / f64 state", "workload": rb["desc"],
                                "n_systems": rb["n_systems"], "n_nucleotides": rb["n_total"],
                                "engine": rb["stats"]["engine"]}
        if not args.no_rna:
            try:
                rr = run_plain("rna", "mixed", args.steps, args.warmup, args.sort_every, args.seed)
                extra["rna"] = {"value": rr["value"], "unit": "nucleotide-steps/s", "ms_per_step": rr["ms_per_step"],
                                "steps": args.steps, "warmup": args.warmup, "dtype": "f64 forces / f64 state",
                                "workload": rr["desc"], "n_systems": rr["n_systems"], "n_nucleotides": rr["n_total"],
                                "roofline": roofline(rr, copy_gbps, profile_counters("rna", "mixed", rr["stats"])),
                                "cpu_baseline": None if args.no_cpu else rna_cpu_baseline(args.seed)}
            except Exception as e:  # the headline stands on its own
                log(f"rna workload failed: {e}")
        if not args.no_served:
            extra["served"] = served_measure(args.jobs, args.step_scale, args.rate_mult, args.budget_gb, args.seed,
                                             args.small_cap, args.quantum_small)
    cpu = None
    if not args.no_cpu and r["ref_keys"] is not None and r["system"] is not None:
        try:
            cpu = cpu_baselines(r["system"], r["ref_keys"], 1, 1 + args.cpu_steps)
        except Exception as e:  # the GPU figure stands on its own
            log(f"cpu baseline failed: {e}")
    print(json.dumps(line(r, args, roof, cpu, extra)), flush=True)


if __name__ == "__main__":
    main()
