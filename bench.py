#!/usr/bin/env python3
"""Benchmark of the DVM time-step hot path on one MI355X.

Workload (one "step"): the isolated regularised Biot-Savart direct sum —
N = 262,144 blobs (x ~ U[-1,15], z ~ U[-3,3], circ ~ N(0, 2e-4) mean-removed,
sigma = 0.01) evaluated blob-on-blob (computeSelfInduced) plus 1,024 probes on
z = 0 (computeInducedAt).  Inputs are resident in HBM before timing starts.

value = pair-interactions/s of the FP32-compensated form (the chosen form);
the plain FP64 form, its error against FP64, and the reference-semantics CPU
oracle (OpenMP, this host's cores) are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 11  # 2 sub, 3 for r^2, 1 reciprocal, 1 for K*circ, 2 FMAs (4 flop)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (AMD published)
PEAK_FP64_TFLOPS = 78.6


def make_inputs(n, n_probes, seed=2024):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 15.0, n)
    z = rng.uniform(-3.0, 3.0, n)
    c = rng.normal(0.0, 2e-4, n)
    c -= c.mean()
    s = np.full(n, 0.01)
    tx = np.linspace(-1.0, 15.0, n_probes)
    tz = np.zeros(n_probes)
    return x, z, c, s, tx, tz


def time_form(ctx, dvm, dev, prec, steps, warmup, outs):
    import torch
    X, Z, C, S, TX, TZ = dev
    (u, w), (pu, pw) = outs

    variants = {}

    def step():
        dvm.bs_self(ctx, X, Z, C, S, prec=prec, out=(u, w))
        variants["self"] = ctx.f32c_variant()
        dvm.bs_targets(ctx, X, Z, C, S, TX, TZ, prec=prec, out=(pu, pw))
        variants["targets"] = ctx.f32c_variant()

    # at least one untimed pass: under measured tuning it picks the variants
    for _ in range(max(warmup, 1)):
        step()
    torch.cuda.synchronize()
    ctx.prof_reset()
    ctx.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.prof_enable(False)
    name = "f32c" if prec == dvm.F32C else "f64"
    self_ms, self_n = ctx.prof_read("bs_self_" + name)
    tgt_ms, tgt_n = ctx.prof_read("bs_targets_" + name)
    return elapsed, (self_ms / max(self_n, 1)), (tgt_ms / max(tgt_n, 1)), variants


def host_cpu():
    """CPU model, sockets x cores per socket x threads per core (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    sockets = int(info.get("Socket(s)", "0") or 0)
    cores = int(info.get("Core(s) per socket", "0") or 0)
    return {"model": info.get("Model name", "unknown"), "sockets": sockets,
            "cores_per_socket": cores, "physical_cores": sockets * cores,
            "threads_per_core": int(info.get("Thread(s) per core", "0") or 0),
            "logical_cpus": os.cpu_count()}


def cpu_baseline(x, z, c, s, budget_s):
    """The oracle (reference loop order, FP64, OpenMP) on a bounded sample of
    targets; each target sums over all N sources with the self pair skipped.
    Timed with a build of the reference's flags for this host (-O3
    -march=native, src/CMakeLists.txt:43), made here before timing; the
    checker's portable build (contraction off) is the fallback."""
    import subprocess
    import tempfile

    import oracle
    path = os.path.join(tempfile.gettempdir(), f"dvmpp_oracle_native_{os.getpid()}.so")
    build = "-O3 -march=native -fopenmp (the reference's flags, built on this host)"
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                        f"NATIVE_OUT={path}"], check=True, timeout=300,
                       capture_output=True)
        using = oracle.native_lib(path)
    except (OSError, subprocess.SubprocessError, RuntimeError) as e:
        log(f"cpu baseline: native build failed ({e}); timing the checker's build")
        using = oracle.lib()
        build = "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (the checker's build)"
    n = x.size
    order = np.random.default_rng(1).permutation(n)
    chunk = 4096
    done = 0
    t0 = time.perf_counter()
    while done < n and time.perf_counter() - t0 < budget_s:
        idx = order[done:done + chunk]
        oracle.bs_targets(x, z, c, s, x[idx], z[idx], using=using)
        done += idx.size
    dt = time.perf_counter() - t0
    try:
        os.remove(path)
    except OSError:
        pass
    return {"value": done * n / dt, "unit": "pair-interactions/s",
            "cores": int(using.orc_num_threads()), "kind": "port",
            "sample": f"{done} blob targets x {n} sources (self pair skipped), FP64 reference "
                      f"loop order, OpenMP, {dt:.1f} s",
            "build": build, "host_cpu": host_cpu()}


def fmm_rate(ctx, dvm, dev, u64, w64, pairs_self, reps=10):
    """The blob-on-blob sum of the same blobs by the fast multipole form at
    two orders: time per sum, error against the FP64 direct result, and the
    direct-sum pair rate that time corresponds to."""
    import torch
    out = {}
    for order, theta in ((10, 0.5), (20, 0.5)):
        ctx.set_fmm(order, theta)
        u, w = dvm.bs_self(ctx, *dev[:4], prec=dvm.FMM)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            u, w = dvm.bs_self(ctx, *dev[:4], prec=dvm.FMM)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        uu, ww = u.cpu().numpy(), w.cpu().numpy()
        err = max(np.abs(uu - u64).max() / np.abs(u64).max(), np.abs(ww - w64).max() / np.abs(w64).max())
        out[f"order{order}_theta{theta}"] = {"ms": ms, "normwise_max_err_vs_f64": float(err),
                                             "leaf_level": ctx.fmm_last_level(),
                                             "equivalent_pair_interactions_per_s":
                                                 pairs_self / (ms * 1e-3)}
    log("fmm: " + ", ".join(f"{k} {v['ms']:.2f} ms err {v['normwise_max_err_vs_f64']:.1e}"
                            for k, v in out.items()))
    return out


def host_aos_rate(ctx, dvm, x, z, c, s):
    """Blob-on-blob sum through the host-pointer entry point (AoS particles in
    host memory, results back in host memory): the PCIe-inclusive rate."""
    import ctypes
    n = x.size
    parts = np.ascontiguousarray(np.stack([x, z, c, s], axis=1))
    res = np.empty((n, 2))
    lib = dvm.load()
    args = (ctx.h, dvm.F32C, n, parts.ctypes.data_as(ctypes.c_void_p),
            res.ctypes.data_as(ctypes.c_void_p))
    dvm.check(ctx.h, lib.dvm_bs_self_host(*args))
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        dvm.check(ctx.h, lib.dvm_bs_self_host(*args))
    dt = (time.perf_counter() - t0) / reps
    return {"pair_interactions_per_s": n * (n - 1) / dt, "ms_per_call": dt * 1e3,
            "note": "dvm_bs_self_host: pageable host AoS in and out, synchronous"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def plain_cylinder_phase(ctx, steps, cpu_steps):
    """Plain cylinder (256 panels, Euler, dt 0.005, REFLECT, seed 0): steps/s of
    the FP64 parity form (reference random stream), of the chosen form (FP32C
    sums, device random stream), of the fast multipole form (order 16,
    device stream), and of the CPU oracle over the first `cpu_steps` steps
    (same window timed on the GPU)."""
    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.workloads import plain_cylinder
    px, pz, sp = plain_cylinder()
    out = {"workload": f"circular cylinder r=0.5, 256 panels, Re=1000, Euler dt=0.005, REFLECT, "
                       f"{steps} steps from rest"}

    def run(prec, rng):
        p = dsim.SimParams(sp, dsim.EULER, dsim.REFLECT, 0, prec, rng, dsim.remesh_params())
        s = dsim.Sim(ctx, px, pz, p)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(cpu_steps)
        torch.cuda.synchronize()
        t_win = time.perf_counter() - t0
        s.run(steps - cpu_steps)
        torch.cuda.synchronize()
        t_all = time.perf_counter() - t0
        n, cd = s.size(), s.loads()[2]
        s.close()
        return t_win, t_all, n, cd

    w64, a64, n64, cd64 = run(dvm.F64, dsim.RNG_REFERENCE)
    log(f"plain cylinder FP64: {steps} steps in {a64:.2f} s, N = {n64}")
    w32, a32, n32, cd32 = run(dvm.F32C, dsim.RNG_DEVICE)
    log(f"plain cylinder FP32C: {steps} steps in {a32:.2f} s, N = {n32}")
    out["f64_reference_stream"] = {"steps_per_s": steps / a64, "final_blobs": n64, "cd": cd64,
                                   "window_steps_per_s": cpu_steps / w64}
    out["f32c_device_stream"] = {"steps_per_s": steps / a32, "final_blobs": n32, "cd": cd32,
                                 "window_steps_per_s": cpu_steps / w32}
    ctx.set_fmm(16, 0.5)
    wf, af, nf, cdf = run(dvm.FMM, dsim.RNG_DEVICE)
    log(f"plain cylinder FMM: {steps} steps in {af:.2f} s, N = {nf}")
    out["fmm_order16_device_stream"] = {"steps_per_s": steps / af, "final_blobs": nf, "cd": cdf,
                                        "window_steps_per_s": cpu_steps / wf}

    # The fast forms against FP64 from the same state (device stream, keyed by
    # step and blob, so both draw the same increments while their blob counts
    # agree): the developed wake of `steps` FP64 steps is copied into the fast
    # form once, then both run k_steps free (no re-sync), compared each step.
    def lockstep(prec, k_steps=10):
        mk = lambda pr: dsim.Sim(ctx, px, pz, dsim.SimParams(sp, dsim.EULER, dsim.REFLECT, 0, pr,
                                                              dsim.RNG_DEVICE,
                                                              dsim.remesh_params()))
        a, b = mk(dvm.F64), mk(prec)
        a.run(steps)
        b.run(steps)
        b.set_state(a.get_state())
        dcd = dcl = dv = 0.0
        for _ in range(k_steps):
            a.step()
            b.step()
            sa, sb = a.get_state(), b.get_state()
            if sa["x"].size == sb["x"].size:
                for f in ("u", "w"):
                    dv = max(dv, float(np.abs(sb[f] - sa[f]).max() / np.abs(sa[f]).max()))
            la, lb = a.loads(), b.loads()
            dcd, dcl = max(dcd, abs(lb[2] - la[2])), max(dcl, abs(lb[3] - la[3]))
        n = a.size()
        a.close()
        b.close()
        return {"steps": k_steps, "blobs": n, "max_abs_dcd": dcd, "max_abs_dcl": dcl,
                "max_normwise_velocity_err": dv}

    out["fast_forms_vs_f64_from_same_state"] = {"f32c": lockstep(dvm.F32C),
                                                 "fmm_order16": lockstep(dvm.FMM)}
    log("fast forms vs FP64 from the same state: "
        + json.dumps(out["fast_forms_vs_f64_from_same_state"]))
    import oracle
    sim = oracle.Sim(px, pz, oracle.make_params(rho=sp.rho, nu=sp.nu, dt=sp.dt, ux=sp.ux,
                                                uz=sp.uz, cutoff_exp=sp.cutoff_exp,
                                                max_npv=sp.max_npv, scheme=0, crossing=2, seed=0))
    t0 = time.perf_counter()
    for _ in range(cpu_steps):
        sim.step()
    tc = time.perf_counter() - t0
    out["cpu_oracle_window"] = {"steps": cpu_steps, "steps_per_s": cpu_steps / tc,
                                "final_blobs": sim.size(),
                                "cores": int(oracle.lib().orc_num_threads())}
    log(f"CPU oracle: {cpu_steps} steps in {tc:.2f} s")
    return out


def c4_remesh():
    """The C3 remesh settings the ensemble and served workloads run with (M4'
    every 20 steps on a 0.01 grid over [-5, 25] x [-10, 10], core size 0.0321),
    which hold each wake to the 1k-400k blobs of the ensemble configuration."""
    from dvmpp_amd import sim as dsim
    return dsim.remesh_params(frequency=20, kernel=dsim.M4PRIME, xmin=-5.0, xmax=25.0,
                              zmin=-10.0, zmax=10.0, dx=0.01, dz=0.01, sigma=0.0321,
                              circ_threshold=1e-10)


def ensemble_phase(ctx, count, window, plain_count, ref_count, seed=3, fmm_min=0):
    """The 128-body ensemble (steps ~ U{200..2000}) in one batch (FP32C,
    graph-replayed steps, remesh), every body run to its own step count or to
    `window` steps if that is smaller (window 0: the full counts).  With
    fmm_min > 0 the batch runs a second time with slots of at least fmm_min
    blobs taking the fast multipole sum (dvm_batch_set_fmm, order 12, theta
    0.5).  Beside it the plain FP64 form one simulation at a time, on a
    random subset of the bodies over the same steps, with the device random
    stream and with the reference stream; the ratio is batched aggregate
    steps/s over plain aggregate steps/s of that subset."""
    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.batch import Batch, batch_config
    from dvmpp_amd.workloads import ensemble
    rp = c4_remesh()
    jobs = ensemble(count)
    run = [min(st, window) if window else st for *_, st in jobs]
    cap = 1 << 17 if 0 < window <= 300 else 1 << 20
    total = sum(run)

    def batched(fmm):
        b = Batch(ctx, batch_config(slots=count, slot_capacity=cap, max_panels=512,
                                    prec=dvm.F32C, use_graph=True, remesh=rp))
        t0 = time.perf_counter()
        slots = b.admit_many([(px, pz, sp, sd, st) for (px, pz, sp, sd, _), st in zip(jobs, run)])
        if fmm:
            b.set_fmm(fmm, order=12, theta=0.5)
        torch.cuda.synchronize()
        ta = time.perf_counter() - t0
        status = {}
        pairs = 0.0
        chunk = 50
        t0 = time.perf_counter()
        done_steps = 0
        while len(status) < count:
            b.step(chunk)
            done_steps += chunk
            # sizes after the chunk: a coarse sum of N^2 over the chunk's steps
            for s in slots:
                if s not in status:
                    pairs += float(b.info(s)["n"]) ** 2 * chunk
            status.update(dict(b.poll(count)))
            if done_steps % 250 == 0:
                log(f"ensemble batch{' (FMM slots)' if fmm else ''}: step {done_steps}, "
                    f"{count - len(status)} running, {time.perf_counter() - t0:.1f} s")
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        blobs = [b.info(s)["n"] for s in slots]
        ok = sum(1 for st in status.values() if st == 0)
        b.close()
        log(f"ensemble batch{' (FMM slots)' if fmm else ''}: {count} bodies, {total} steps in "
            f"{tb:.2f} s (admission {ta:.2f} s)")
        return {"aggregate_steps_per_s": total / tb, "seconds": tb, "admission_seconds": ta,
                "completed_ok": ok, "slot_capacity": cap,
                "final_blobs_min_median_max": [int(min(blobs)), int(np.median(blobs)),
                                               int(max(blobs))],
                "pair_interactions_per_s_estimate": pairs / tb}

    res = batched(0)
    tb, ta = res["seconds"], res["admission_seconds"]
    res_fmm = batched(fmm_min) if fmm_min else None
    pick = np.random.default_rng(seed).choice(count, max(plain_count, ref_count), replace=False)
    plain = {}
    for name, mode, m in (("device_stream", dsim.RNG_DEVICE, plain_count),
                          ("reference_stream", dsim.RNG_REFERENCE, ref_count)):
        if m == 0:
            continue
        t_set = t_run = 0.0
        steps = 0
        for k in pick[:m]:
            px, pz, sp, sd, _ = jobs[k]
            t1 = time.perf_counter()
            sim = dsim.Sim(ctx, px, pz, dsim.SimParams(sp, dsim.EULER, dsim.REFLECT, sd, dvm.F64,
                                                        mode, rp))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            sim.run(run[k])
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            sim.close()
            t_set += t2 - t1
            t_run += t3 - t2
            steps += run[k]
            log(f"ensemble plain FP64 ({name}): body {k}, {run[k]} steps in {t3 - t2:.2f} s")
        plain[name] = {"bodies": int(m), "steps": steps, "seconds": t_run, "setup_seconds": t_set,
                       "aggregate_steps_per_s": steps / t_run,
                       "speedup_of_batched": (total / tb) / (steps / t_run),
                       "speedup_of_batched_with_setup": (total / (tb + ta)) /
                                                        (steps / (t_run + t_set))}
        log(f"ensemble plain FP64 ({name}): {m} bodies, {steps} steps in {t_run:.2f} s")
    win = f"the first {window} steps of each" if window else "each body's full step count"
    return {"workload": f"{count} bodies (ellipses AR~U[1,4], AoA~U[0,30] deg; rounded squares), "
                        f"128-512 panels, nu log-U[2e-4,2e-3], dt 0.005, steps ~ U{{200..2000}} "
                        f"({win}: {total} simulation steps), Euler, REFLECT, M4' remesh every 20 "
                        f"steps at 0.01 (sigma 0.0321)",
            "batched_f32c_graph": res,
            **({"batched_fmm_slots": dict(res_fmm, fmm_min_blobs=fmm_min,
                                          form="slots of at least fmm_min_blobs blobs: fast "
                                               "multipole sum (order 12, theta 0.5); the others "
                                               "FP32C direct",
                                          speedup_over_plain={
                                              k: res_fmm["aggregate_steps_per_s"]
                                              / v["aggregate_steps_per_s"]
                                              for k, v in plain.items()})}
               if res_fmm else {}),
            "plain_f64_per_simulation": plain,
            "plain_subset": f"{len(pick)} bodies drawn at random (seed {seed}): "
                            f"{sorted(int(k) for k in pick)}"}


def served_phase(ctx, count, rate, steps_range, slots, fmm_min=0):
    """Poisson arrivals at `rate`/s of ensemble-distribution requests (128-512
    panels, step counts ~ U{steps_range}, remesh), served by continuous
    batching (FP32C, graph-replayed steps) with capacity classes of 128K, 512K
    and 1M blobs per slot allocated within a 200 GB HBM budget; with
    fmm_min > 0, slots of at least fmm_min blobs take the fast multipole sum
    (order 12, theta 0.5)."""
    import dvmpp_amd as dvm
    from dvmpp_amd.batch import batch_config
    from dvmpp_amd.serve import Request, Server
    from dvmpp_amd.workloads import request_trace
    trace = request_trace(count, rate=rate, steps_range=steps_range)
    reqs = [Request(i, t, px, pz, sp, seed, st) for i, (t, (px, pz, sp, seed, st)) in enumerate(trace)]
    srv = Server(ctx, batch_config(slots=slots, slot_capacity=1 << 20, max_panels=512,
                                   prec=dvm.F32C, use_graph=True, remesh=c4_remesh()),
                 budget_bytes=200 << 30, classes=[1 << 17, 1 << 19, 1 << 20], fmm_min=fmm_min)
    done, wall = srv.run(reqs)
    peak = srv.peak_bytes
    srv.close()
    ok = [r for r in done if r.status == 0]
    lat = np.array([r.latency for r in ok]) if ok else np.array([np.nan])
    steps = sum(r.steps for r in ok)
    log(f"served: {len(done)} requests in {wall:.2f} s")
    fm = (f", slots of at least {fmm_min} blobs by the fast multipole sum (order 12, theta 0.5)"
          if fmm_min else "")
    return {"workload": f"{count} requests, Poisson {rate}/s, ensemble bodies (128-512 panels), "
                        f"steps ~ U{{{steps_range[0]}..{steps_range[1]}}}, M4' remesh every 20 "
                        f"steps at 0.01, {slots} slots per batch region, capacity classes "
                        f"128K/512K/1M blobs, 200 GB budget{fm}",
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "aggregate_steps_per_s": steps / wall, "wall_s": wall,
            "completed_ok": len(ok), "rejected": sum(1 for r in done if r.status == 2),
            "peak_allocated_bytes": peak,
            "requests_by_class": {str(c): sum(1 for r in done if r.capacity == c)
                                  for c in (1 << 17, 1 << 19, 1 << 20)}}


def rk3_plateau_phase(ctx, steps, prec=None, fmm=None):
    """RK3 cylinder (512 panels, nu 5e-4, dt 0.0025, remesh M4' every 20
    steps) timed per step at a fixed-N snapshot of the plateau: a remeshed
    wake of ~1.5 M blobs is loaded into the simulation, one step runs untimed,
    then `steps` full RK3 steps (3 blob-on-blob sums, 3 sheet solves and
    sheet-on-blob sums, diffusion, release, crossing) are timed; one remesh of
    the same state is timed separately.  `prec` is the blob-on-blob sum form
    (FP32C by default); `fmm` = (order, theta) for the FMM form."""
    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.workloads import rk3_cylinder, wake_snapshot
    px, pz, sp, sig = rk3_cylinder()
    rm = dsim.remesh_params(frequency=20, kernel=dsim.M4PRIME, xmin=-5.0, xmax=25.0, zmin=-10.0,
                            zmax=10.0, dx=0.01, dz=0.01, sigma=sig, circ_threshold=1e-10)
    snap = wake_snapshot(sigma=sig)
    n0 = snap["x"].size
    prec = dvm.F32C if prec is None else prec
    if fmm is not None:
        ctx.set_fmm(*fmm)
    p = dsim.SimParams(sp, dsim.RK3, dsim.REFLECT, 0, prec, dsim.RNG_DEVICE, rm)
    s = dsim.Sim(ctx, px, pz, p)
    s.set_state(snap)
    s.run(1)
    torch.cuda.synchronize()
    n1 = s.size()
    t0 = time.perf_counter()
    s.run(steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    n2 = s.size()
    s.close()
    log(f"RK3 plateau: {steps} steps at N = {n1}..{n2} in {dt * steps:.2f} s")
    X, Z, C, S = (torch.from_numpy(snap[k]).cuda() for k in ("x", "z", "circ", "sigma"))
    dsim.remesh(ctx, rm, X, Z, C, S)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = dsim.remesh(ctx, rm, X, Z, C, S)
    torch.cuda.synchronize()
    tr = time.perf_counter() - t0
    nmean = 0.5 * (n1 + n2)
    form = ("FP32C sums" if prec == dvm.F32C else
            f"FMM sums (order {fmm[0]}, theta {fmm[1]})" if prec == dvm.FMM else "FP64 sums")
    return {"workload": f"RK3 cylinder, 512 panels, nu 5e-4, dt 0.0025, REFLECT, {form}, "
                        f"device random stream; remeshed-wake snapshot of {n0} blobs",
            "seconds_per_step": dt, "steps_timed": steps, "blobs": [n1, n2],
            "steps_per_s": 1.0 / dt,
            "pair_interactions_per_s_in_step": 3.0 * nmean * (nmean - 1) / dt,
            "remesh_ms": tr * 1e3, "remesh_blobs_out": int(out[0].numel())}


def graph_sim_phase(ctx, window=40):
    """Single simulations in the chosen form with graph-replayed steps: each
    runs as a one-slot batch (FP32C sums, device random stream, counts kept on
    the device, each step two replayed hipGraphs with the RK3 stages inside
    the first).  (1) The RK3 cylinder at its plateau: the 1.49M-blob
    remeshed wake loaded into the slot, one step untimed, then two timed.
    (2) The RK3 cylinder run from rest to step 4000 (M4' remesh every 20
    steps).  (3) The plain cylinder's 400 Euler steps from rest, its first
    `window` steps also timed alone."""
    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.batch import Batch, batch_config
    from dvmpp_amd.workloads import plain_cylinder, rk3_cylinder, wake_snapshot
    out = {}
    px, pz, sp, sig = rk3_cylinder()
    rm = dsim.remesh_params(frequency=20, kernel=dsim.M4PRIME, xmin=-5.0, xmax=25.0, zmin=-10.0,
                            zmax=10.0, dx=0.01, dz=0.01, sigma=sig, circ_threshold=1e-10)

    def one_slot(scheme, cap, remesh, body, params, steps):
        b = Batch(ctx, batch_config(slots=1, slot_capacity=cap, max_panels=512, scheme=scheme,
                                    crossing=dsim.REFLECT, prec=dvm.F32C, use_graph=True,
                                    remesh=remesh))
        return b, b.admit(body[0], body[1], params, 0, steps)

    snap = wake_snapshot(sigma=sig)
    b, slot = one_slot(dsim.RK3, 1 << 21, rm, (px, pz), sp, 1 << 30)
    b.set_state(slot, snap["x"], snap["z"], snap["circ"], snap["sigma"])
    b.step(1)
    torch.cuda.synchronize()
    n1 = b.info(slot)["n"]
    t0 = time.perf_counter()
    b.step(2)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 2
    n2 = b.info(slot)["n"]
    b.close()
    nmean = 0.5 * (n1 + n2)
    out["rk3_plateau"] = {"blobs": [int(n1), int(n2)], "seconds_per_step": dt,
                          "steps_per_s": 1.0 / dt,
                          "pair_interactions_per_s_in_step": 3.0 * nmean * (nmean - 1) / dt}
    log(f"graph RK3 plateau: 2 steps at N = {n1}..{n2} in {2 * dt:.2f} s")

    run_steps = 4000
    b, slot = one_slot(dsim.RK3, 1 << 20, rm, (px, pz), sp, run_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trace = []
    for k in range(0, run_steps, 500):
        b.step(500)
        info = b.info(slot)
        trace.append([k + 500, int(info["n"]), round(time.perf_counter() - t0, 3)])
        log(f"graph RK3 run: step {k + 500}, N = {info['n']}")
    torch.cuda.synchronize()
    tr = time.perf_counter() - t0
    info = b.info(slot)
    b.close()
    out["rk3_run_4000"] = {"seconds": tr, "steps_per_s": run_steps / tr,
                           "final_blobs": int(info["n"]), "status": int(info["status"]),
                           "final_cd_cl": [info["loads"][2], info["loads"][3]],
                           "trace_step_blobs_seconds": trace}

    cpx, cpz, csp = plain_cylinder()
    b, slot = one_slot(dsim.EULER, 1 << 18, dsim.remesh_params(), (cpx, cpz), csp, 400)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(window)
    torch.cuda.synchronize()
    tw = time.perf_counter() - t0
    b.step(400 - window)
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    info = b.info(slot)
    b.close()
    out["plain_cylinder_400"] = {"seconds": tc, "steps_per_s": 400 / tc,
                                 "window_steps_per_s": window / tw,
                                 "final_blobs": int(info["n"]), "cd": info["loads"][2]}
    log(f"graph plain cylinder: 400 steps in {tc:.2f} s, N = {info['n']}")
    out["workload"] = ("one-slot batches, FP32C sums, device random stream, graph-replayed "
                       "steps: RK3 cylinder (512 panels, nu 5e-4, dt 0.0025, M4' remesh every "
                       "20 steps at 0.01) at its 1.49M-blob plateau and from rest to step 4000; "
                       "plain cylinder (256 panels, Euler, dt 0.005) for 400 steps")
    return out


def rk3_run_phase(ctx, steps, fmm=(12, 0.5)):
    """The RK3 cylinder run itself: from rest, `steps` RK3 steps (512
    panels, nu 5e-4, dt 0.0025, M4' remesh every 20 steps at dx = 0.01),
    blob sums by the fast multipole form, device random stream; wall time of
    the whole run, with the blob count and loads along the way."""
    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.workloads import rk3_cylinder
    px, pz, sp, sig = rk3_cylinder()
    rm = dsim.remesh_params(frequency=20, kernel=dsim.M4PRIME, xmin=-5.0, xmax=25.0, zmin=-10.0,
                            zmax=10.0, dx=0.01, dz=0.01, sigma=sig, circ_threshold=1e-10)
    ctx.set_fmm(*fmm)
    s = dsim.Sim(ctx, px, pz, dsim.SimParams(sp, dsim.RK3, dsim.REFLECT, 0, dvm.FMM,
                                             dsim.RNG_DEVICE, rm))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trace = []
    done = 0
    while done < steps:
        k = min(500, steps - done)
        s.run(k)
        done += k
        trace.append([done, s.size(), round(time.perf_counter() - t0, 3)])
        log(f"RK3 run: step {done}, N = {s.size()}, {trace[-1][2]:.1f} s")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cd, cl = s.loads()[2:]
    s.close()
    return {"workload": f"RK3 cylinder from rest, {steps} steps, 512 panels, nu 5e-4, dt 0.0025, "
                        f"REFLECT, M4' remesh every 20 steps at dx 0.01, FMM sums (order {fmm[0]}, "
                        f"theta {fmm[1]}), device random stream",
            "steps": steps, "seconds": wall, "steps_per_s": steps / wall,
            "final_blobs": trace[-1][1], "trace_step_blobs_seconds": trace,
            "final_cd_cl": [cd, cl]}


def merge_phase(ctx, steps, thresholds=(1e-8, 1e-7)):
    """merge_blobs on a grown wake: the C1 cylinder (256 panels, FP32C sums,
    device stream) run `steps` steps, then one merge pass per threshold on a
    copy of its blobs, timed, with the merged fraction and the work counts."""
    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.bodies import cylinder
    px, pz = cylinder(256)
    s = dsim.Sim(ctx, px, pz, dsim.sim_params(seed=3, prec=dvm.F32C, rng_mode=dsim.RNG_DEVICE))
    s.run(steps)
    st = s.get_state()
    s.close()
    n = st["x"].size
    fields = ("x", "z", "circ", "sigma", "u", "w", "uvs", "wvs")
    out = {"workload": f"circular cylinder r=0.5, 256 panels, Re=1000, {steps} Euler steps, "
                       f"FP32C sums, device stream: {n} blobs", "blobs": n}
    lib = dvm._lib.load()
    import ctypes
    for thr in thresholds:
        d = [torch.from_numpy(st[k]).cuda() for k in fields]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k = dsim.merge_blobs(ctx, thr, *d)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        a, b = ctypes.c_int64(), ctypes.c_int64()
        lib.dvm_merge_stats(ctx.h, ctypes.byref(a), ctypes.byref(b))
        out[f"threshold_{thr:g}"] = {"ms": ms, "merged": k, "merged_fraction": k / n,
                                     "candidate_blobs": a.value, "recomputed_sequences": b.value}
        log(f"merge threshold {thr:g}: {k} merges ({100 * k / n:.1f} %) in {ms:.1f} ms")
    return out


def bandwidth_phase(ctx, reps=20):
    """The HBM-bound blob operations at the RK3 plateau size: the remeshed
    wake of ~1.5 M blobs (workloads.wake_snapshot) with its last 1 % of blobs
    moved inside the 512-panel body (as blobs released last and diffused
    through the surface would be) so that the crossing's compaction runs.
    Single-kernel entry points are timed over `reps` back-to-back calls
    between HIP events on the context's stream; the multi-kernel ones
    (delete_inside: flags, scan, compaction; remesh: sort, gather, scan, emit,
    with a host read of the output size) per call with a synchronisation.
    GB/s = algorithmic bytes (FP64 SoA fields read + written, per blob) /
    time, against the 8 TB/s HBM peak."""
    import ctypes

    import torch

    import dvmpp_amd as dvm
    from dvmpp_amd import sim as dsim
    from dvmpp_amd.workloads import rk3_cylinder, wake_snapshot
    lib = dvm._lib.load()
    px, pz, sp, sig = rk3_cylinder()
    snap = wake_snapshot(sigma=sig)
    n = snap["x"].size
    rng = np.random.default_rng(11)
    k_in = n // 100
    idx = np.arange(n - k_in, n)
    rr = 0.45 * np.sqrt(rng.uniform(0.0, 1.0, k_in))
    th = rng.uniform(-np.pi, np.pi, k_in)
    x0, z0 = snap["x"].copy(), snap["z"].copy()
    x0[idx], z0[idx] = rr * np.cos(th), rr * np.sin(th)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X0, Z0 = dev(x0), dev(z0)
    C0, S0 = dev(snap["circ"]), dev(snap["sigma"])
    vel = [dev(rng.normal(0.0, 0.1, n)) for _ in range(4)]  # u w uvs wvs
    q1, q2 = dev(np.zeros(n)), dev(np.zeros(n))
    tape = dev(rng.normal(0.0, 1e-3, 2 * n))
    ids = torch.arange(n, dtype=torch.int32, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    sheet = dsim.Sheet(ctx, px, pz, sp)
    flags = torch.empty(n, dtype=torch.int32, device="cuda")
    X, Z = X0.clone(), Z0.clone()
    u, w, uvs, wvs = vel

    def events(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    chk = lambda st: dvm._lib.check(ctx.h, st)
    out = {"workload": f"remeshed RK3-cylinder wake, {n} blobs ({k_in} inside the 512-panel "
                       f"body), FP64 SoA resident in HBM", "blobs": n, "peak_GBps": 8000.0}

    def rec(name, ms, bytes_per_blob, extra_bytes=0):
        gbs = (bytes_per_blob * n + extra_bytes) / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "bytes_per_blob": bytes_per_blob, "GBps": gbs,
                     "frac_of_hbm_peak": gbs / 8000.0}
        log(f"bandwidth: {name} {ms * 1e3:.1f} us, {gbs:.0f} GB/s")

    rec("advect_euler", events(lambda: chk(lib.dvm_advect_euler(
        ctx.h, n, 1.0, 0.0, 0.0025, P(u), P(w), P(uvs), P(wvs), P(X), P(Z)))), 64)
    rec("advect_rk3_stage", events(lambda: chk(lib.dvm_advect_rk3_stage(
        ctx.h, n, -17.0 / 32.0, 8.0 / 9.0, 1.0, 0.0, 0.0025, P(u), P(w), P(uvs), P(wvs), P(q1),
        P(q2), P(X), P(Z)))), 96)
    rec("diffuse_tape", events(lambda: chk(lib.dvm_diffuse(ctx.h, n, P(tape), P(X), P(Z)))), 48)
    rec("crossing_flags", events(lambda: chk(lib.dvm_sheet_inside(sheet.h, n, P(X0), P(Z0),
                                                                  P(flags)))), 20)
    inside = int(flags.sum().item())

    # delete_inside: in place, so each call starts from a fresh copy (not timed)
    fields0 = [X0, Z0, C0, S0] + vel
    ms = []
    for _ in range(5):
        f = [t.clone() for t in fields0]
        idc = ids.clone()
        m = ctypes.c_int64()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        chk(lib.dvm_sheet_delete_inside(sheet.h, n, *[P(t) for t in f], P(idc), ctypes.byref(m)))
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    # flags (x, z in, flag out) + stable compaction of 8 FP64 fields and the ID
    rec("delete_inside", float(np.median(ms)), 20 + 136)
    out["delete_inside"]["deleted"] = n - m.value
    out["crossing_flags"]["inside"] = inside

    rm = dsim.remesh_params(frequency=20, kernel=dsim.M4PRIME, xmin=-5.0, xmax=25.0, zmin=-10.0,
                            zmax=10.0, dx=0.01, dz=0.01, sigma=sig, circ_threshold=1e-10)
    # one dvm_remesh call into outputs sized beforehand (the call counts the
    # output, reads the count to the host, then emits)
    n_new = dsim.remesh(ctx, rm, X0, Z0, C0, S0)[0].numel()
    o = [torch.empty(n_new, dtype=torch.float64, device="cuda") for _ in range(4)]
    m = ctypes.c_int64()
    ms = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        chk(lib.dvm_remesh(ctx.h, ctypes.byref(rm), n, P(X0), P(Z0), P(C0), P(S0), n_new,
                           *[P(t) for t in o], ctypes.byref(m)))
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    nodes = 3001 * 2001
    rec("remesh_m4prime", float(np.median(ms)), 24, 16 * nodes + 32 * n_new)
    out["remesh_m4prime"].update({"mesh_nodes": nodes, "blobs_out": n_new,
                                  "bytes": "24 N + 16 nodes + 32 N_out"})
    sheet.close()
    return out


def load_traffic(kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--probes", type=int, default=1024)
    ap.add_argument("--f64-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", choices=("measure", "table"), default="table",
                    help="F32C launch variant: the built-in size table (the library default) "
                         "or timed on first use per size (measure)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (reference kernel "
                         "for scaling HBM byte counters)")
    ap.add_argument("--phases", default="sum,cylinder,ensemble,served,rk3,graph,merge,bandwidth",
                    help="comma list: sum (isolated sum, always run), cylinder, ensemble, "
                         "served, rk3, rk3run, merge, bandwidth; ensemble_full (every body to "
                         "its own step count, several minutes) only when named")
    ap.add_argument("--rk3-steps", type=int, default=2)
    ap.add_argument("--rk3-fmm-steps", type=int, default=10)
    ap.add_argument("--rk3-run-steps", type=int, default=4000)
    ap.add_argument("--cylinder-steps", type=int, default=400)
    ap.add_argument("--cylinder-cpu-steps", type=int, default=40)
    ap.add_argument("--ensemble-count", type=int, default=128)
    ap.add_argument("--ensemble-window", type=int, default=300,
                    help="steps per body in the ensemble phase (0: each body's full count)")
    ap.add_argument("--ensemble-plain", type=int, default=8)
    ap.add_argument("--ensemble-plain-ref", type=int, default=8)
    ap.add_argument("--ensemble-full-plain", type=int, default=8,
                    help="ensemble_full phase: bodies of the plain FP64 subset (device stream)")
    ap.add_argument("--ensemble-full-plain-ref", type=int, default=4,
                    help="ensemble_full phase: bodies of the plain FP64 subset (reference stream)")
    ap.add_argument("--ensemble-fmm-min", type=int, default=65536,
                    help="ensemble_full phase: slots of at least this many blobs take the fast "
                         "multipole sum in the second batched run (0: no second run)")
    ap.add_argument("--served-count", type=int, default=8)
    ap.add_argument("--served-steps", default="500,3000")
    ap.add_argument("--served-fmm-min", type=int, default=65536,
                    help="a second served run in which slots of at least this many blobs take "
                         "the fast multipole sum (0: none)")
    args = ap.parse_args()
    phases = set(args.phases.split(","))

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(an ensemble would shard by simulation across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import dvmpp_amd as dvm

    n, npr = args.n, args.probes
    x, z, c, s, tx, tz = make_inputs(n, npr)
    ctx = dvm.Context(0)
    dev = [torch.from_numpy(a).to("cuda") for a in (x, z, c, s, tx, tz)]

    def outs(m):
        return (torch.empty(m, dtype=torch.float64, device="cuda"),
                torch.empty(m, dtype=torch.float64, device="cuda"))

    o32 = (outs(n), outs(npr))
    o64 = (outs(n), outs(npr))
    pairs_self = n * (n - 1)
    pairs_probe = npr * n
    pairs = pairs_self + pairs_probe

    ctx.set_tuning(dvm.TUNE_MEASURE if args.tune == "measure" else dvm.TUNE_TABLE)
    el32, self32_ms, tgt32_ms, variants = time_form(ctx, dvm, dev, dvm.F32C, args.steps,
                                                    args.warmup, o32)
    el64, self64_ms, tgt64_ms, _ = time_form(ctx, dvm, dev, dvm.F64, args.f64_steps, 1, o64)

    # accuracy of the timed FP32C result against the FP64 form
    u32, w32 = (t.cpu().numpy() for t in o32[0])
    u64, w64 = (t.cpu().numpy() for t in o64[0])
    v = np.hypot(u64, w64)
    d = np.hypot(u32 - u64, w32 - w64)
    rel = d / np.maximum(v, 1e-300)

    ms_step = el32 / args.steps * 1e3
    value = pairs / (el32 / args.steps)
    achieved_tf = FLOP_PER_PAIR * pairs_self / (self32_ms * 1e-3) / 1e12
    line = {
        "metric": "Biot-Savart pair-interactions/s + sim steps/s at 1 GPU; p99 served latency",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"isolated Biot-Savart direct sum: {n} blobs self + {npr} probes, "
                        "x~U[-1,15], z~U[-3,3], circ~N(0,2e-4) mean-removed, sigma=0.01",
            "n_blobs": n, "n_probes": npr, "pairs_per_step": pairs,
            "form": "FP32 pair terms, tile-recentred Morton-sorted sources, Kahan across tiles",
            "f32c_variant": {k: dict(zip(("tile", "block", "share", "occupancy"), v))
                             for k, v in variants.items()},
            "tuning": args.tune,
        },
        "roofline": {
            "bound": "valu",
            "kernel": "bs_f32c_kernel (blob-on-blob)",
            "achieved": achieved_tf,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tf / PEAK_FP32_TFLOPS,
            "traffic": load_traffic("bs_f32c_kernel"),
            "kernel_ms": self32_ms,
            "flop_per_launch": FLOP_PER_PAIR * pairs_self,
        },
        "f32c_error_vs_f64": {"normwise_max": float(d.max() / v.max()),
                              "p99_rel": float(np.percentile(rel, 99)),
                              "p50_rel": float(np.percentile(rel, 50))},
        "f64_plain": {
            "pair_interactions_per_s": pairs / (el64 / args.f64_steps),
            "ms_per_step": el64 / args.f64_steps * 1e3,
            "self_kernel_ms": self64_ms,
            "achieved_tflops": FLOP_PER_PAIR * pairs_self / (self64_ms * 1e-3) / 1e12,
            "frac_of_fp64_peak": FLOP_PER_PAIR * pairs_self / (self64_ms * 1e-3) / 1e12
            / PEAK_FP64_TFLOPS,
        },
        "probe_kernel_ms": {"f32c": tgt32_ms, "f64": tgt64_ms},
    }
    line["fmm_self"] = fmm_rate(ctx, dvm, dev, u64, w64, pairs_self)
    line["host_aos_f32c"] = host_aos_rate(ctx, dvm, x, z, c, s)
    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = nbytes
    log(f"isolated sum: FP32C {ms_step:.2f} ms/step, FP64 {el64 / args.f64_steps * 1e3:.2f} ms/step")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(x, z, c, s, args.cpu_seconds)
        log("cpu baseline done")
    del dev, o32, o64
    torch.cuda.empty_cache()
    # The simulation phases grow N every step: they use the size table, so no
    # variant timing runs inside their timed windows.
    ctx.set_tuning(dvm.TUNE_TABLE)
    if "cylinder" in phases:
        line["sim_steps_plain_cylinder"] = plain_cylinder_phase(ctx, args.cylinder_steps,
                                                                args.cylinder_cpu_steps)
    if "ensemble" in phases:
        line["ensemble"] = ensemble_phase(ctx, args.ensemble_count, args.ensemble_window,
                                          args.ensemble_plain, args.ensemble_plain_ref)
    if "ensemble_full" in phases:  # several minutes: not in the default list
        line["ensemble_full"] = ensemble_phase(ctx, args.ensemble_count, 0,
                                               args.ensemble_full_plain,
                                               args.ensemble_full_plain_ref,
                                               fmm_min=args.ensemble_fmm_min)
    if "served" in phases:
        lo, hi = (int(v) for v in args.served_steps.split(","))
        line["served"] = served_phase(ctx, args.served_count, 20.0, (lo, hi), slots=32)
        if args.served_fmm_min:
            line["served_fmm_slots"] = served_phase(ctx, args.served_count, 20.0, (lo, hi),
                                                    slots=32, fmm_min=args.served_fmm_min)
    if "bandwidth" in phases:
        line["bandwidth_kernels"] = bandwidth_phase(ctx)
    if "merge" in phases:
        line["merge"] = merge_phase(ctx, args.cylinder_steps)
    if "rk3run" in phases:
        line["rk3_cylinder_run_fmm"] = rk3_run_phase(ctx, args.rk3_run_steps)
    if "graph" in phases:
        line["graph_single_simulation"] = graph_sim_phase(ctx, args.cylinder_cpu_steps)
    if "rk3" in phases:
        line["rk3_cylinder_plateau"] = rk3_plateau_phase(ctx, args.rk3_steps)
        line["rk3_cylinder_plateau_fmm"] = rk3_plateau_phase(ctx, args.rk3_fmm_steps, dvm.FMM,
                                                             (12, 0.5))
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
