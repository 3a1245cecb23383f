"""Benchmark of the LBM time step on one MI355X.

Workload (the headline configuration): D3Q19 TRT (Lambda 3/16, tau 0.8),
192^3 periodic, one squirmer of radius 12 (B1 = 1e-3, beta from --beta,
position and orientation uniform from the seed), Ladd moving-wall bounce-back
with momentum-exchange force/torque, populations stored in 16 bits (the
custom 1-4-11 float over (f - w rho0)/w, "f16c") with FP32 collision.  One
step = one tick! of the reference.

Prints ONE JSON line: MLUPS (all nodes x steps / wall second on the default
run path, hipGraph replays of the one-launch tick), the roofline of the step
kernel (algorithmic bytes 2*Q*sizeof(storage) per node over its average
duration from HIP events on its stream, and over the wall-clock step), the
CPU baseline (the oracle, a C restatement of the reference step, on this
host's cores), FP64/FP32/IEEE-FP16 variants of the same step, the errors of
the FP32, FP16 and f16c storages against FP64 after --accuracy-steps steps
for beta in {-5, 0, 5} (the velocity field outside the sphere is the graded
figure), and the two smaller configurations (D2Q9 squirmer 128^2 FP64,
D3Q19 channel 256x64x64).

The "batch" object measures the batched workload: 64 independent D3Q19 TRT
periodic boxes, edges drawn from {48, 64, 96}, one squirmer each (radius
6-10, beta uniform in [-5, 5], B1 = 1e-3), FP32 storage, all advanced 2000
ticks together (one launch per tick over every box, graph-replayed), with
each box's velocity field and swimming speed against the same box run alone
in FP64.

The "adjoint" object measures the gradient workload: D3Q19 TRT 96^3
periodic, one squirmer of radius 10 (B1 = 1e-3, beta = 0), a Guo force field
drawn N(0, 1e-7), FP32 forward and adjoint over 2000 ticks with checkpoints
every 100; loss = (mean speed over the last 500 ticks - 1e-4)^2 + sum F^2;
gradients with respect to B1, B2 and the force field, checked against FP64
central finite differences (steps 1e-6 of each parameter's scale) along 16
seeded directions.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "lbmengine.jl_amd"))

METRIC = "MLUPS + % HBM peak, D3Q19 TRT squirmer 192^3; FP16-store vs FP64 at err≤1e-3"
HBM_PEAK_GBS = 8000.0
BYTES = {"f64": 8, "f32": 4, "f16": 2, "f16c": 2}
ACC_BOUND = 1e-3  # velocity-field error against FP64 the headline is quoted at


def workload_kwargs(storage: str, n: int):
    return dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                kineticParameter=3 / 16, xlims=(-n // 2, n // 2 - 1), storage=storage)


def squirmer_kwargs(seed: int, n: int, r: float, beta: float):
    rng = np.random.default_rng(seed)
    lo, hi = -n // 2 + r + 6, n // 2 - 1 - r - 6
    pos = rng.uniform(lo, hi, 3)
    e = rng.standard_normal(3)
    return dict(radius=r, position=pos, swimmingSpeed=2e-3 / 3, beta=beta, swimmingDirection=e / np.linalg.norm(e),
                coupleForces=True, coupleTorques=True)


def build_model(lh, storage, args):
    m = lh.modelInit(**workload_kwargs(storage, args.n))
    lh.addSquirmer(m, **squirmer_kwargs(args.seed, args.n, args.radius, args.beta))
    return m


def run_wall(m, steps, warmup):
    """Wall time of `steps` ticks on the default path (hipGraph replays of
    the step kernel), after `warmup` ticks, synchronised on both sides."""
    m.run(warmup)
    m.sync()
    t0 = time.perf_counter()
    m.run(steps)
    m.sync()
    return time.perf_counter() - t0


def batch_specs(seed, count=64):
    """The batched workload's boxes: (edge, radius, beta, position, axis) each."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n = int(rng.choice([48, 64, 96]))
        r = float(rng.uniform(6.0, 10.0))
        beta = float(rng.uniform(-5.0, 5.0))
        lo, hi = -n // 2 + r + 6, n // 2 - 1 - r - 6
        pos = rng.uniform(lo, hi, 3)
        e = rng.standard_normal(3)
        specs.append((n, r, beta, pos, e / np.linalg.norm(e)))
    return specs


def batch_boxes(lh, storage, seed, count=64):
    """The batched workload's boxes (same seed, same boxes for any storage)."""
    models = []
    for n, r, beta, pos, e in batch_specs(seed, count):
        m = lh.modelInit(**workload_kwargs(storage, n))
        lh.addSquirmer(m, radius=r, position=pos, swimmingSpeed=2e-3 / 3, beta=beta, swimmingDirection=e,
                       coupleForces=True, coupleTorques=True)
        models.append(m)
    return models


def batch_bench(lh, args):
    """64 uneven FP32 boxes advanced together; throughput, batched kernel
    roofline, and per-box errors against FP64 runs of each box alone."""
    Q = 19
    boxes = batch_boxes(lh, "f32", args.seed)
    nodes = sum(int(np.prod(m.resolved.n[:3])) for m in boxes)
    b = lh.LBMbatch(boxes)
    if args.batch_block:
        parts = args.batch_block.split(",")
        if parts[0].startswith("r"):
            # row tiles: rTHREADS[,tile order[,low occupancy]]
            v = [int(parts[0][1:])] + [int(x) for x in parts[1:]]
            block = (v[0], 0, 0)
            b.set_row_tiles(v[0])
            v = v[:1] + [0, 0] + v[1:]
        else:
            v = [int(x) for x in parts]
            block = tuple(v[:3])
            b.set_block(*block)
        if len(v) > 3:
            b.set_tile_order(v[3])
        if len(v) > 4:
            b.set_occupancy(v[4])
    else:
        block, tuned_ms = b.autotune(5)
    b.run(args.batch_warmup)
    b.sync()
    t0 = time.perf_counter()
    b.run(args.batch_steps)
    b.sync()
    wall = time.perf_counter() - t0
    b.set_timing(True)
    b.run(200)
    b.sync()
    b.set_timing(False)
    kms, sms = b.last_run_times()
    bytes_per_launch = 2 * Q * BYTES["f32"] * nodes
    res = {"workload": "64 D3Q19 TRT periodic boxes, edges from {48,64,96}, one squirmer each (r 6-10, "
                       "beta U[-5,5], B1=1e-3), f32 storage, one launch per tick over all boxes",
           "boxes": len(boxes), "nodes": nodes, "steps": args.batch_steps, "warmup": args.batch_warmup,
           "MLUPS": round(nodes * args.batch_steps / wall / 1e6, 1),
           "ms_per_step": round(wall / args.batch_steps * 1e3, 5), "block": list(block),
           "row_tiles": b.get_row_tiles(),
           "tile_order": b.get_tile_order(), "low_occupancy": b.get_occupancy(),
           "kernel_ms": round(kms, 5), "step_ms_device": round(sms, 5),
           "hbm_frac": round(bytes_per_launch / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if args.batch_accuracy:
        ticks = args.batch_warmup + args.batch_steps + 200
        ref = batch_boxes(lh, "f64", args.seed)
        eu, es = [], []
        for lo, hi in zip(boxes, ref):
            hi.run(ticks)
            hi.sync()
            u64, v64 = hi.fluidVelocity, hi.particles[0].velocity
            eu.append(float(np.linalg.norm(lo.fluidVelocity - u64) / np.linalg.norm(u64)))
            es.append(float(abs(np.linalg.norm(lo.particles[0].velocity) - np.linalg.norm(v64))
                            / np.linalg.norm(v64)))
            del hi
        res["accuracy_vs_f64"] = {"ticks": ticks, "u_rel_l2_max": max(eu), "u_rel_l2_median": float(np.median(eu)),
                                  "speed_rel_max": max(es)}
    return res


def adjoint_model(lh, storage, n, r, B1, B2, F, pos, e):
    m = lh.modelInit(**workload_kwargs(storage, n), forceDensity=np.asfortranarray(F))
    lh.addSquirmer(m, radius=r, position=pos, swimmingSpeed=2 * B1 / 3, beta=B2 / B1, swimmingDirection=e,
                   coupleForces=True, coupleTorques=True)
    return m


def adjoint_bench(lh, args):
    """The gradient workload: adjoint time and the check against FP64
    central differences."""
    from lbmengine_hip import adjoint
    n, r, ticks, K, window = ADJ_N, ADJ_R, args.adjoint_ticks, 100, min(500, args.adjoint_ticks)
    F, pos, e = adjoint_workload(args.seed)
    rng = np.random.default_rng(args.seed + 6)
    B1, B2 = 1e-3, 0.0
    opts = dict(ticks=ticks, window=window, target_speed=1e-4, force_penalty=1.0)
    m = adjoint_model(lh, "f32", n, r, B1, B2, F, pos, e)
    m.autotune(5)  # the recomputation's forward kernels
    g = adjoint.swimming_loss_gradient(m, checkpoint_every=K, **opts)
    del m
    # the gradient holds the starting populations fixed; so do the runs below
    f0 = adjoint_model(lh, "f64", n, r, B1, B2, F, pos, e).distributions
    res = {"workload": f"D3Q19 TRT {n}^3 periodic, squirmer r={r:g} B1=1e-3 beta=0, Guo force field N(0,1e-7), "
                       f"f32 forward and adjoint, {ticks} ticks, checkpoints every {K}, speed averaged over the "
                       f"last {window}",
           "loss": g.loss, "dL_dB1": float(g.dB1[0]), "dL_dB2": float(g.dB2[0]),
           "dL_dF_norm": float(np.linalg.norm(g.dF)),
           "forward_ms": round(g.forward_ms, 1), "reverse_ms": round(g.reverse_ms, 1),
           "reverse_ms_per_tick": round(g.reverse_ms / ticks, 4)}
    errs, steps, switched = [], [], 0
    for d in range(args.adjoint_dirs):
        d1, d2, dF = rng.standard_normal(), rng.standard_normal(), rng.standard_normal(F.shape)
        h = 1e-6
        while True:
            s1, s2, sF = h * B1 * d1, h * B1 * d2, h * 1e-7 * dF
            ta, tb = [], []
            mp = adjoint_model(lh, "f64", n, r, B1 + s1, B2 + s2, F + sF, pos, e)
            mm = adjoint_model(lh, "f64", n, r, B1 - s1, B2 - s2, F - sF, pos, e)
            mp.distributions = f0
            mm.distributions = f0
            Lp = adjoint.swimming_loss(mp, force_field=F + sF, trajectory=ta, **opts)
            Lm = adjoint.swimming_loss(mm, force_field=F - sF, trajectory=tb, **opts)
            # a link switch between the two runs is a jump of the geometry,
            # not part of the derivative: shrink the step until there is none
            if adjoint.link_switches(mp, ta, tb) == 0 or h < 1e-10:
                break
            switched += 1
            h /= 10.0
        fd = (Lp - Lm) / 2.0
        ad = g.dB1[0] * s1 + g.dB2[0] * s2 + float(np.sum(g.dF * sF))
        errs.append(abs(ad - fd) / abs(fd))
        steps.append(h)
    if errs:
        res["fd_check"] = {"directions": len(errs), "step": "1e-6 of each parameter's scale, divided by 10 "
                                                             "while the two runs' link sets differ at some tick",
                           "steps_shrunk": switched, "smallest_step": min(steps),
                           "max_rel_err": max(errs), "median_rel_err": float(np.median(errs)),
                           "rel_errs": [float(x) for x in errs]}
    return res


def kernel_times(m, steps):
    """Average device time of the step kernel and of a whole tick, from HIP
    events recorded on the kernel's stream around every step kernel."""
    m.set_timing(True)
    m.run(steps)
    m.sync()
    m.set_timing(False)
    return m.last_run_times()


# Bytes one launch of each kernel moves by the algorithm (read + write), for
# the workloads this script runs: the step kernel reads and writes every
# population once; the batched kernel does so for every box; the adjoint's
# bulk and near kernels read each node's moments of the tick (rho and u,
# 16 B in FP32), read and write the adjoint populations (compute type), read
# the force field and update the segment's dL/dF (compute type): 16 + 19*4*2
# + 12 + 24 = 204 B per node in FP32 (the post-collision populations the near
# kernel reads at particle links add well under 0.1 %, the FP64 fold of
# dL/dF once per segment of 100 ticks 0.2 %).
ADJ_N, ADJ_BYTES_PER_NODE, ADJ_R = 96, 204, 10.0


def adjoint_workload(seed):
    """(force field, swimmer position, swimming direction) of the gradient
    workload."""
    n, r = ADJ_N, ADJ_R
    rng = np.random.default_rng(seed + 5)
    F = rng.normal(0.0, 1e-7, (n, n, n, 3))
    lo, hi = -n // 2 + r + 6, n // 2 - 1 - r - 6
    pos = rng.uniform(lo, hi, 3)
    e = rng.standard_normal(3)
    e /= np.linalg.norm(e)
    return F, pos, e


def adjoint_near_nodes(seed):
    """Nodes of the swimmer's near box (its link box grown by one node) at the
    start: the adjoint's near kernel takes these, the bulk kernel the rest."""
    _, pos, _ = adjoint_workload(seed)
    x0, count = -(ADJ_N // 2), 1
    for k in range(3):
        lo = math.floor(pos[k] - ADJ_R - x0) - 3
        hi = math.ceil(pos[k] + ADJ_R - x0) + 3
        count *= min(hi, ADJ_N - 1) - max(lo, 0) + 1
    return count


def _kernel_family(name):
    """Short key of a traced kernel: the fused step kernel of a storage (either
    occupancy build), the fused batched kernel, the adjoint's bulk and near
    kernels, the copy kernel or the pull kernel of an element type."""
    n = name.replace(" ", "")
    if n.startswith("void"):
        n = n[4:]
    if "copy_kernel" in n:
        return "copy_kernel"
    targs = n.split("<", 1)[1].split(">", 1)[0].split(",") if "<" in n else []
    # step_kernel<Q, S, COLL, FM, MODE, OCC>: MODE 1 is the fused kernel
    if n.startswith("lbm::step_kernel<") and len(targs) >= 5 and targs[4] == "1":
        return "step_kernel/" + {"0": "f64", "1": "f32", "2": "f16", "3": "f16c"}[targs[1]]
    if n.startswith("lbm::batch_step_kernel<") and len(targs) >= 5 and targs[4] == "1":
        return "batch_step_kernel"
    if n.startswith("lbm::adjoint_near_kernel<"):
        return "adjoint_near_kernel"
    if n.startswith("lbm::adjoint_kernel<"):
        return "adjoint_kernel"
    if n.startswith("lbm::pull_kernel<") and targs:
        return "pull_kernel/" + targs[0]
    return None


def _pmc_means(run_dir):
    """Mean counter value per kernel family in a rocprofv3 --pmc output folder."""
    import csv
    from collections import defaultdict
    files = sorted(Path(run_dir).glob("**/*counter_collection.csv"))
    if not files:
        raise SystemExit(f"no counter_collection.csv under {run_dir}")
    acc = defaultdict(list)
    for f in files:
        for r in csv.DictReader(open(f)):
            fam = _kernel_family(r["Kernel_Name"])
            if fam:
                acc[(fam, r["Counter_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}


def traffic_summary(fetch_dir, write_dir, args, argv):
    """HBM bytes per launch of every traced kernel from two rocprofv3 --pmc
    runs of this script (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch),
    each scaled by what the same counter reads for the copy kernel, whose
    traffic is known (it reads and writes half the headline step's bytes)."""
    f, fc = _pmc_means(fetch_dir)
    w, wc = _pmc_means(write_dir)
    copy_bytes = 2 * 19 * BYTES[args.storage] * args.n ** 3 // 2 // 16 * 16
    kf = copy_bytes / (f[("copy_kernel", "FETCH_SIZE")] * 1024)
    kw = copy_bytes / (w[("copy_kernel", "WRITE_SIZE")] * 1024)
    algo = {"batch_step_kernel": 2 * 19 * 4 * sum(s[0] ** 3 for s in batch_specs(args.seed)),
            "adjoint_kernel": ADJ_BYTES_PER_NODE * (ADJ_N ** 3 - adjoint_near_nodes(args.seed)),
            "adjoint_near_kernel": ADJ_BYTES_PER_NODE * adjoint_near_nodes(args.seed)}
    # (only the headline storage's step kernel: the others in a run belong
    # to other workloads, such as the adjoint's 96^3 forward steps)
    algo[f"step_kernel/{args.storage}"] = 2 * 19 * BYTES[args.storage] * args.n ** 3
    # the headline's pull (the step's gathers and stores alone, same bytes)
    pull_t = {2: "unsignedshort", 4: "unsignedint", 8: "unsignedlong"}[BYTES[args.storage]]
    algo[f"pull_kernel/{pull_t}"] = 2 * 19 * BYTES[args.storage] * args.n ** 3
    made_by = [f"rocprofv3 --pmc FETCH_SIZE -d {fetch_dir} -o pmc --output-format csv -- python3 bench.py {argv}",
               f"rocprofv3 --pmc WRITE_SIZE -d {write_dir} -o pmc --output-format csv -- python3 bench.py {argv}",
               f"python3 bench.py --traffic-from {fetch_dir} {write_dir} {argv}"]
    out = {}
    for fam, a in algo.items():
        if (fam, "FETCH_SIZE") not in f or (fam, "WRITE_SIZE") not in w:
            continue
        rd = f[(fam, "FETCH_SIZE")] * 1024 * kf
        wr = w[(fam, "WRITE_SIZE")] * 1024 * kw
        out[fam] = {"kernel": fam, "hbm_bytes_per_launch": round(rd + wr), "read_bytes": round(rd),
                    "write_bytes": round(wr), "algorithmic_bytes_per_launch": a,
                    "traffic_over_algorithmic": round((rd + wr) / a, 4),
                    "fetch_scale": round(kf, 4), "write_scale": round(kw, 4),
                    "raw_kib": {"fetch": f[(fam, "FETCH_SIZE")], "write": w[(fam, "WRITE_SIZE")],
                                "copy_fetch": f[("copy_kernel", "FETCH_SIZE")],
                                "copy_write": w[("copy_kernel", "WRITE_SIZE")]},
                    "dispatches": {"fetch": fc[(fam, "FETCH_SIZE")], "write": wc[(fam, "WRITE_SIZE")]},
                    "made_by": made_by}
    return out


def cpu_baseline(args):
    """The oracle (FP64 C restatement of the reference step, OpenMP over
    nodes) on the same workload, timed for a bounded number of steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import harness
    kw = workload_kwargs("f64", args.n)
    o = harness.oracle_model(**kw)
    o.add("squirmer", **squirmer_kwargs(args.seed, args.n, args.radius, args.beta))
    o.run(1)  # first touch
    steps, t = 0, 0.0
    t0 = time.perf_counter()
    while steps < args.cpu_max_steps and t < args.cpu_seconds:
        o.run(1)
        steps += 1
        t = time.perf_counter() - t0
    nodes = args.n ** 3
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": nodes * steps / t / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "sample": f"{steps} steps of the same 192^3 squirmer workload in FP64 ({t:.1f} s), "
                      f"oracle/lbm_oracle.c with OpenMP over nodes ({cores} threads)"}


def small_configs(lh, args):
    """The two smaller configurations: MLUPS on the run path and the step
    kernel's time from HIP events (with its algorithmic-byte rate)."""
    out = {}
    # D2Q9 BGK 128^2 periodic FP64, squirmer disk r = 8 at the centre, random
    # orientation from the seed, CPU-backend semantics, 1000 steps
    rng = np.random.default_rng(args.seed)
    th = rng.uniform(0, 2 * np.pi)
    m = lh.modelInit(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8,
                     collisionModel="bgk", xlims=(-64, 63), storage="f64", semantics="cpu")
    lh.addSquirmer(m, radius=8.0, swimmingSpeed=2e-3 / 3, beta=0.0, swimmingDirection=[np.cos(th), np.sin(th)],
                   coupleForces=True, coupleTorques=True)
    m.autotune(20)
    w = run_wall(m, 1000, 100)
    k, s = kernel_times(m, 1000)
    n1 = 128 * 128
    out["C1"] = {"workload": "D2Q9 BGK tau=0.8 128^2 periodic, squirmer disk r=8 B1=1e-3 beta=0, CPU-backend "
                             "semantics, f64 storage, 1000 steps",
                 "MLUPS": round(n1 * 1000 / w / 1e6, 1), "ms_per_step": round(w / 1000 * 1e3, 5),
                 "kernel_ms": round(k, 5), "step_ms_device": round(s, 5), "block": list(m.get_block()),
                 "hbm_frac_algorithmic": round(2 * 9 * 8 * n1 / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "note": "16,384 nodes (2.4 MB) per launch: launch-latency bound, the working set lives in cache"}
    del m
    # D3Q19 TRT (Lambda 3/16) channel 256 x 64 x 64, walls at z = 0 and 63,
    # uniform Guo force 1e-6 along x, tau 0.9
    for st in ("f32", "f64"):
        m = lh.modelInit(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.9,
                         kineticParameter=3 / 16, xlims=(0, 255), ylims=(0, 63), zlims=(0, 63), walledDimensions=[3],
                         forceDensity=[1e-6, 0, 0], storage=st)
        m.autotune(20)
        w = run_wall(m, 2000, 200)
        k, s = kernel_times(m, 1000)
        n2 = 256 * 64 * 64
        out[f"C2_{st}"] = {"workload": f"D3Q19 TRT tau=0.9 channel 256x64x64, walls at z=0,63, Guo force 1e-6, "
                                       f"{st} storage",
                           "MLUPS": round(n2 * 2000 / w / 1e6, 1), "ms_per_step": round(w / 2000 * 1e3, 5),
                           "kernel_ms": round(k, 5), "step_ms_device": round(s, 5), "block": list(m.get_block()),
                           "hbm_frac_algorithmic": round(2 * 19 * BYTES[st] * n2 / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "note": f"{2 * 19 * BYTES[st] * n2 / 1e6:.0f} MB per launch"
                                   + (": within the 256 MB last-level cache, so the algorithmic rate can exceed "
                                      "what HBM delivers" if 2 * 19 * BYTES[st] * n2 <= 256e6 else
                                      ": more than the 256 MB last-level cache holds")}
        del m
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--storage", default="f16c", choices=list(BYTES))
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--radius", type=float, default=12.0)
    ap.add_argument("--beta", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--accuracy-steps", type=int, default=5000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-max-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--batch-warmup", type=int, default=100)
    ap.add_argument("--batch-accuracy", type=int, default=1, help="0: skip the per-box FP64 comparison")
    ap.add_argument("--batch-block", default="", help="bx,by,bz[,tile order[,low occupancy]] of the batched kernel, or "
                                                      "rTHREADS[,tile order[,low occupancy]] for row tiles, instead of "
                                                      "autotuning")
    ap.add_argument("--no-configs", action="store_true", help="skip the D2Q9 and channel configurations")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-ticks", type=int, default=2000)
    ap.add_argument("--adjoint-dirs", type=int, default=16)
    ap.add_argument("--block", default="", help="bx,by,bz[,tile order[,low occupancy]] of the step kernel instead of autotuning")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs of this script into the traffic file and exit")
    args = ap.parse_args()
    if args.traffic_from:
        # the remaining arguments are those of the profiled runs
        argv = " ".join(a for a in sys.argv[1:] if a not in ("--traffic-from", *args.traffic_from))
        out = traffic_summary(*args.traffic_from, args, argv)
        (ROOT / "profiles").mkdir(exist_ok=True)
        names = {"batch_step_kernel": "traffic_batch_64_boxes.json", "adjoint_kernel": "traffic_adjoint_96.json",
                 "adjoint_near_kernel": "traffic_adjoint_near_96.json"}
        for fam, rec in out.items():
            if fam.startswith("pull_kernel/"):
                names[fam] = f"traffic_pull_{args.storage}_{args.n}.json"
            dst = ROOT / "profiles" / (names[fam] if fam in names else f"traffic_step_{fam.split('/')[1]}_{args.n}.json")
            dst.write_text(json.dumps(rec, indent=1) + "\n")
        print(json.dumps(out))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import lbmengine_hip as lh
    nodes = args.n ** 3
    Q = 19

    m = build_model(lh, args.storage, args)
    if args.block:
        v = [int(x) for x in args.block.split(",")]
        m.set_block(*v[:3])
        if len(v) > 3:
            m.set_tile_order(v[3])
        if len(v) > 4:
            m.set_occupancy(v[4])
        tuned_ms = float("nan")
    else:
        block, tuned_ms = m.autotune(20)
    wall = run_wall(m, args.steps, args.warmup)
    mlups = nodes * args.steps / wall / 1e6
    # the roofline's kernel time: HIP events around every step kernel
    kms, sms = kernel_times(m, min(args.steps, 1000))
    bytes_per_launch = 2 * Q * BYTES[args.storage] * nodes
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_step_{args.storage}_{args.n}.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
    # achievable ceiling: a plain device copy moving the same bytes per launch
    copy_ms = lh._capi.measure_copy_ms(bytes_per_launch // 2, 10)
    copy_gbs = bytes_per_launch / (copy_ms * 1e-3) / 1e9
    # the ceiling of the step's access pattern: its gathers and stores alone
    pull_ms = lh._capi.measure_pull_ms(lh._capi.STORAGE[args.storage], args.n, args.n, args.n, 10)
    pull_gbs = bytes_per_launch / (pull_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.storage,
        "dtype_note": {"f16c": "16-bit float (1 sign, 4 exponent, 11 mantissa bits) over (f - w rho0)/w, "
                               "FP32 collision", "f16": "IEEE FP16 over (f - w rho0) 2^14/w, FP32 collision",
                       "f32": "FP32 deviations", "f64": "FP64 populations"}[args.storage],
        "data": "synthetic (uniform fluid at rest, squirmer position/orientation from numpy seed)",
        "config": {"workload": f"D3Q19 TRT Lambda=3/16 tau=0.8 {args.n}^3 periodic, squirmer r={args.radius:g} "
                               f"B1=1e-3 beta={args.beta:g}, Ladd + momentum exchange, {args.storage} storage",
                   "nodes": nodes, "block": list(m.get_block()), "tile_order": m.get_tile_order(), "low_occupancy": m.get_occupancy(), "autotuned_kernel_ms": None if tuned_ms != tuned_ms else round(tuned_ms, 5)},
        "event_timed_MLUPS": round(nodes / (sms * 1e-3) / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "step_kernel", "kernel_ms": round(kms, 5), "step_ms_device": round(sms, 5),
                     "bytes_per_launch": bytes_per_launch,
                     "frac_wall_clock": round(bytes_per_launch / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                     "copy_ceiling_GBs": round(copy_gbs, 1), "copy_ms": round(copy_ms, 5),
                     "frac_of_copy": round(achieved / copy_gbs, 4),
                     "pull_ceiling_GBs": round(pull_gbs, 1), "pull_ms": round(pull_ms, 5),
                     "frac_of_pull": round(achieved / pull_gbs, 4)},
    }
    del m

    if not args.no_variants:
        var = {}
        for st in ("f64", "f32", "f16", "f16c"):
            if st == args.storage:
                var[st] = {"MLUPS": out["value"], "kernel_ms": out["roofline"]["kernel_ms"],
                           "hbm_frac": out["roofline"]["frac"], "pull_ms": out["roofline"]["pull_ms"],
                           "frac_of_pull": out["roofline"]["frac_of_pull"]}
                continue
            mv = build_model(lh, st, args)
            mv.autotune(10)
            ks = max(args.steps // 4, 100)
            w = run_wall(mv, ks, args.warmup // 4)
            k, _ = kernel_times(mv, min(ks, 500))
            pm = lh._capi.measure_pull_ms(lh._capi.STORAGE[st], args.n, args.n, args.n, 10)
            var[st] = {"MLUPS": round(nodes * ks / w / 1e6, 1), "kernel_ms": round(k, 5),
                       "hbm_frac": round(2 * Q * BYTES[st] * nodes / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "pull_ms": round(pm, 5), "frac_of_pull": round(pm / k, 4),
                       "block": list(mv.get_block()), "tile_order": mv.get_tile_order(),
                       "low_occupancy": mv.get_occupancy()}
            del mv
        var["speedup_f16c_over_f64"] = round(var["f16c"]["MLUPS"] / var["f64"]["MLUPS"], 2)
        var["speedup_f16_over_f64"] = round(var["f16"]["MLUPS"] / var["f64"]["MLUPS"], 2)
        out["variants"] = var

        # the reduced-precision storages against FP64 after the same ticks,
        # for each beta of the headline sweep; the main figure is the
        # velocity field outside the sphere (the reference also streams the
        # fluid inside it)
        acc = {"steps": args.accuracy_steps, "bound": ACC_BOUND, "graded": "u_rel_l2_exterior"}
        worst = {}
        for beta in (-5.0, 0.0, 5.0):
            a2 = argparse.Namespace(**vars(args))
            a2.beta = beta
            ref = build_model(lh, "f64", a2)
            ref.run(args.accuracy_steps)
            ref.sync()
            u64 = ref.fluidVelocity
            v64 = ref.particles[0].velocity
            R = ref.particles[0].position
            c = [np.asarray(x) for x in ref.resolved.coords]
            d2 = ((c[0][:, None, None] - R[0]) ** 2 + (c[1][None, :, None] - R[1]) ** 2
                  + (c[2][None, None, :] - R[2]) ** 2)
            ext = d2 >= args.radius ** 2
            del ref
            row = {"swimming_speed_f64": float(np.linalg.norm(v64))}
            for st in ("f32", "f16", "f16c"):
                mm = build_model(lh, st, a2)
                mm.run(args.accuracy_steps)
                mm.sync()
                u = mm.fluidVelocity
                v = mm.particles[0].velocity
                r = {"u_rel_l2_exterior": float(np.linalg.norm((u - u64)[ext]) / np.linalg.norm(u64[ext])),
                     "u_rel_l2": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                     "speed_rel": float(abs(np.linalg.norm(v) - np.linalg.norm(v64)) / np.linalg.norm(v64)),
                     "velocity_rel": float(np.linalg.norm(v - v64) / np.linalg.norm(v64))}
                row[st] = r
                worst[st] = max(worst.get(st, 0.0), r["u_rel_l2_exterior"], r["speed_rel"])
                del mm
            acc[f"beta={beta:g}"] = row
        acc["worst_over_beta"] = worst
        acc["within_bound"] = {st: bool(e <= ACC_BOUND) for st, e in worst.items()}
        out["accuracy_vs_f64"] = acc

    if not args.no_configs:
        out["configs"] = small_configs(lh, args)
    if not args.no_batch:
        out["batch"] = batch_bench(lh, args)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(lh, args)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
