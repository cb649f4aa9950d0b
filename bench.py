"""Throughput benchmark of the MI355X test-particle tracer.

Workload (N=1): the synthetic beam pipe -- straight cylinder R 4 cm, L 2 m,
~100k triangles, 8 side ports (R 2 cm, 20 cm) ending in pump disks of
sticking 0.05, 6 NEG bands of sticking 0.01, outgassing uniform per wall
area, 256x256 textures on 16 facets and 90x180 angle maps on 4 facets.
A step is one launch that desorbs `--des-per-step` new particles; particles
still in flight when a launch's queue runs dry stay on the device and the
next launch resumes them first (the reference's particles likewise persist
between SimulationMCStep calls). The timed job is K steps followed by one
drain launch that traces every remaining particle to its end, so every hit
of the K*des-per-step desorptions is inside the timed region; the default
100 steps of 1e6 is the full 1e8-desorption job.

Prints ONE JSON line: hits/s (reference definition: bounces + absorptions +
volume scatters per wall second) over K timed steps, with the roofline
figure of the trace kernel (algorithmic bytes per hit from the reference-
equivalent BVH, HIP-event kernel time) and the CPU oracle timed on this
host's cores on a bounded sample of the same workload. The same line
carries the other legs of the headline metric (skipped with --no-extra):
  tube          the 1e7-desorption tube job (R 1 cm, L 10 cm, 64 x 50 wall
                quads) and its own CPU baseline;
  batched       64 uneven beam-pipe variants (a quarter of the 256 of the
                batched workload, 2e3-2e4 desorptions each: a tenth) in
                one launch, captured in a hipGraph, against one launch per
                variant;
  adjoint       cost of a run that also accumulates the 32 x 64 sticking
                Jacobian against the same forward run without it;
  matched_sigma time the GPU and the CPU oracle need to bring the standard
                error of one tally (the fraction of particles the 8 pumps
                absorb) to the same value.
Counter-derived figures (HBM traffic, atomics, VALU issue) come from the
rocprofv3 summaries under profiles/, and only when the summary records the
hash of the kernel sources this build was made from.

--workload tube / batched / adjoint run one leg alone (batched at 256
variants of 2e4-2e5 desorptions; adjoint with its finite-difference check).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per hit (one wall event), counted on the reference-
# equivalent BVH (binary SAH, <= 2 facets per leaf):
#   B_hit = 32 N_node + 48 N_prim + 96 + 16 N_field
# 32 B per visited node (FP32 box + child word), 48 B per tested triangle
# (plane, origin and edge vectors as FP32 equivalents), 96 B of the hit
# facet's parameters and 16 B (read + write of 8 B) per tally field updated:
# 7 counter fields per hit plus 3 texture fields per hit on a textured facet.
NODE_BYTES, PRIM_BYTES, FACET_BYTES, FIELD_BYTES, FIELDS_PER_HIT, TEX_FIELDS = 32, 48, 96, 16, 7, 3


def wall_rays(n_rays, seed=5, radius=4.0, length=200.0):
    """Particle rays of the beam pipe: points on the main wall with cosine
    (Lambertian) directions about the inward normal, i.e. diffuse bounces."""
    import numpy as np
    rng = np.random.default_rng(seed)
    z = rng.uniform(1, length - 1, n_rays)
    a = rng.uniform(0, 2 * math.pi, n_rays)
    r = radius * (1 - 1e-4)
    o = np.stack([r * np.cos(a), r * np.sin(a), z], 1)
    nrm = -np.stack([np.cos(a), np.sin(a), np.zeros(n_rays)], 1)
    t1 = np.stack([-np.sin(a), np.cos(a), np.zeros(n_rays)], 1)
    t2 = np.cross(nrm, t1)
    u1, u2 = rng.random(n_rays), rng.random(n_rays)
    ct, st, ph = np.sqrt(1 - u1), np.sqrt(u1), 2 * math.pi * u2
    d = nrm * ct[:, None] + t1 * (st * np.cos(ph))[:, None] + t2 * (st * np.sin(ph))[:, None]
    return o, d


def algorithmic_bytes_per_hit(model, tex_hit_fraction, n_rays=20000, radius=4.0, length=200.0):
    """Mean BVH nodes and facet tests per ray on the oracle's BVH, over
    diffuse-bounce rays of the workload (main cylinder of the given radius
    and length)."""
    from oracle.mfcpu import CpuModel
    cm = CpuModel(model)
    o, d = wall_rays(n_rays, radius=radius, length=length)
    nodes, prims = cm.bvh_stats(o, d)
    fields = FIELDS_PER_HIT + TEX_FIELDS * tex_hit_fraction
    b = NODE_BYTES * nodes + PRIM_BYTES * prims + FACET_BYTES + FIELD_BYTES * fields
    return b, nodes, prims


def cpu_baseline(model, run_s=3.0, repeats=5, name="beam-pipe"):
    """The CPU oracle (one tracer per OpenMP thread, thread-local tallies,
    the reference's CPU parallelism) on a bounded sample of the workload:
    a calibration run of at least a second sizes `repeats` runs of about
    `run_s` seconds each (disjoint particle ranges); the median rate is
    reported."""
    from oracle.mfcpu import CpuModel
    cm = CpuModel(model)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS", threads))
    n, first = 2000, 0
    while True:
        t0 = time.perf_counter()
        cm.run(n, seed=1, first=first, threads=threads)
        dt = time.perf_counter() - t0
        first += n
        if dt >= 1.0 or n >= 4_000_000:
            break
        n *= 4
    n_run = int(max(2000, min(20_000_000, n * run_s / max(dt, 1e-3))))
    rates, total_hits, total_s = [], 0, 0.0
    for _ in range(repeats):
        t0 = time.perf_counter()
        t = cm.run(n_run, seed=2, first=first, threads=threads)
        dt = time.perf_counter() - t0
        first += n_run
        hits = int(t.glob[0]["nb_mc_hit"])
        rates.append(hits / dt)
        total_hits += hits
        total_s += dt
    rates.sort()
    return {"value": rates[len(rates) // 2], "unit": "hits/s", "cores": threads, "kind": "port",
            "desorptions_per_s": n_run * repeats / total_s,
            "sample": f"median of {repeats} runs of {n_run} desorptions of the same {name} model "
                      f"({total_hits} hits in {total_s:.1f} s in all), oracle/mfcpu.cpp",
            "runs": rates}


KERNEL_SOURCES = ("molflow_sources_amd/csrc/mfg_trace.hip", "molflow_sources_amd/csrc/mfg_internal.h",
                  "include/molflow_gpu.h")


def code_only(text):
    """C/C++ source without comments, trailing blanks and empty lines (string
    and character literals kept as written)."""
    out, i, n, quote = [], 0, len(text), None
    while i < n:
        c = text[i]
        if quote:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if c == quote:
                quote = None
            i += 1
        elif c in "\"'":
            quote = c
            out.append(c)
            i += 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        else:
            out.append(c)
            i += 1
    return "\n".join(l.rstrip() for l in "".join(out).splitlines() if l.strip())


def kernel_source_hash():
    """sha256 of the code of the kernel sources (the trace kernel and the
    layouts it reads; comments and blank lines left out): a profile summary
    applies to this build only if it records it."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), encoding="utf-8") as fh:
            h.update(code_only(fh.read()).encode())
    return h.hexdigest()


def profile_value(name, key):
    """`key: value` from profiles/<name> when the summary's
    kernel_source_sha256 line names these kernel sources, else None."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    vals = {}
    for line in open(p):
        if ":" in line and not line.startswith((" ", "#")):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if vals.get("kernel_source_sha256") != kernel_source_hash() or key not in vals:
        return None
    try:
        return float(vals[key])
    except ValueError:
        return None


def atomic_peak(device=0, n_addr=1 << 20, per_thread=64):
    """Device-scope FP64 and uint64 atomic-add rates of this GPU on 1M
    distinct counters (molflow_sources_amd/libmfg_microbench.so, a
    measurement helper): the peak the trace kernel's tally atomics are
    reported against."""
    import ctypes
    path = os.path.join(ROOT, "molflow_sources_amd", "libmfg_microbench.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.mfgb_atomic_rates.restype = ctypes.c_int
    lib.mfgb_atomic_rates.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    f, u = ctypes.c_double(), ctypes.c_double()
    if lib.mfgb_atomic_rates(device, n_addr, per_thread, ctypes.byref(f), ctypes.byref(u)) != 0:
        return None
    return {"f64_per_s": f.value, "u64_per_s": u.value, "counters": n_addr,
            "ops": "one FP64 (or uint64) atomic add without return per lane, 8 blocks of 256 per CU"}


def gather_rates(device=0, sizes_mib=(64, 1024), per_thread=64):
    """Rates at which this GPU reads whole 128 B lines at random indices of
    an array of each size (libmfg_microbench.so): the leaf-record and node
    access pattern of the trace kernel. 64 MiB is about the beam pipe's
    working set (Infinity Cache resident), 1 GiB comes from HBM."""
    import ctypes
    path = os.path.join(ROOT, "molflow_sources_amd", "libmfg_microbench.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    if not hasattr(lib, "mfgb_gather_rate"):
        return None
    lib.mfgb_gather_rate.restype = ctypes.c_int
    lib.mfgb_gather_rate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    out = {}
    for mib in sizes_mib:
        r = ctypes.c_double()
        if lib.mfgb_gather_rate(device, mib << 20, per_thread, ctypes.byref(r)) != 0:
            return None
        out[f"{mib}MiB_GBps"] = r.value / 1e9
    return out


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed launches (default: the whole job, 100 for plain, 10 for tube)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--des-per-step", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-extra", action="store_true",
                    help="plain workload only: skip the tube, batched, adjoint and matched-sigma legs")
    ap.add_argument("--no-carry", action="store_true", help="trace every launch's particles to their end")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch autotuner")
    ap.add_argument("--waves", type=int, default=0, help="occupancy bound of the trace kernel (with --no-tune)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this size first (HBM counter calibration)")
    ap.add_argument("--workload", choices=("plain", "tube", "batched", "adjoint"), default="plain",
                    help="plain: the beam-pipe job (default); tube: the 1e7-desorption tube job; batched: "
                         "many uneven variants in one launch vs one launch per variant; adjoint: sticking "
                         "sensitivities vs a forward run")
    ap.add_argument("--variants", type=int, default=256)
    ap.add_argument("--variant-des", type=str, default="2e4,2e5",
                    help="desorptions per variant, log-uniform in [lo,hi]")
    ap.add_argument("--adjoint-des", type=int, default=200_000)
    ap.add_argument("--fd-des", type=int, default=40_000)
    return ap.parse_args(argv)


def main():
    args = parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: runs across several GPUs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        lo, hi = (float(x) for x in args.variant_des.split(","))
        print(json.dumps(batched_leg(args.variants, lo, hi)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(adjoint_leg(args.adjoint_des, args.fd_des)), flush=True)
        return
    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    tube = args.workload == "tube"
    out, info = plain_leg(args, tube)
    if not tube and not args.no_extra:
        progress("tube leg")
        out["tube"] = tube_leg(args)
        progress("batched leg")
        out["batched"] = batched_leg(64, 2e3, 2e4, graph=True)
        progress("adjoint leg")
        out["adjoint"] = adjoint_leg(100_000, 0)
        if out["cpu_baseline"] and "pump_fraction_batches" in info:
            out["matched_sigma"] = matched_sigma(info, out["cpu_baseline"])
    print(json.dumps(out), flush=True)


def plain_leg(args, tube):
    """The timed job on the beam pipe (or the tube): returns the JSON line's
    fields and what matched_sigma needs."""
    from molflow_sources_amd import geometry
    from molflow_sources_amd.simulation import GpuSimulation

    steps = args.steps if args.steps is not None else (10 if tube else 100)
    if tube:
        # R 1 cm, L 10 cm, 64 x 50 wall quads, inlet disk desorbing (cosine),
        # outlet disk of sticking 1, walls of sticking 0: 1e7 desorptions
        model = geometry.tube_model()
        ray_geom = {"radius": 0.998, "length": 10.0}   # just inside the 64-gon's facet planes
    else:
        model = geometry.beam_pipe_model()
        ray_geom = {"radius": 4.0, "length": 200.0}
    sim = GpuSimulation(0)
    sim.load(model)
    sim.rebuild_accel()
    sim.set_launch(args.block, args.blocks_per_cu)
    sim.set_carry(not args.no_carry)
    sim.reset(seed=42)
    tune = None
    if not args.no_tune:
        # occupancy bound and block size for this model (the tuner discards
        # its probe runs: it zeroes the tallies afterwards)
        tr = sim.autotune(args.des_per_step)
        tune = {"waves_per_simd": tr.waves, "block": tr.block, "blocks_per_cu": tr.blocks_per_cu,
                "bvh_width": tr.bvh_width, "leaf_size": tr.leaf_size, "hot_facets": tr.n_hot, "probe_hits_per_s": tr.hits_per_s}
    elif args.waves:
        sim.set_kernel_variant(args.waves)
    max_events = 50_000_000
    for _ in range(args.warmup):
        sim.run(args.des_per_step, max_events=max_events)
    sim.drain(max_events=max_events)
    sim.set_carry(not args.no_carry)
    sim.reset(seed=42)
    hits = 0
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        r = sim.run(args.des_per_step, max_events=max_events)   # synchronous: returns after the device is done
        hits += r.hits
        kernel_ms += r.kernel_ms
    td = time.perf_counter()
    r = sim.drain(max_events=max_events)
    hits += r.hits
    kernel_ms += r.kernel_ms
    wall = time.perf_counter() - t0
    drain_s = time.perf_counter() - td
    t = sim.read()
    assert t.glob[0]["nb_desorbed"] == steps * args.des_per_step and sim.in_flight() == 0
    assert t.glob[0]["nb_mc_hit"] == hits
    value = hits / wall
    info = {"des_per_s": steps * args.des_per_step / wall}
    if not tube and not args.no_extra:
        # batch means of the pump-absorbed fraction over 10 untimed batches
        # (each traced to its end), for the matched-sigma leg
        pumps = list(model.pump_ids)
        fr = []
        sim.set_carry(False)
        for b in range(10):
            sim.clear_tallies()
            sim.reset(seed=1000 + b)
            sim.run(args.des_per_step, max_events=max_events)
            tb = sim.read()
            fr.append(float(tb.facets[0]["nb_abs_equiv"][pumps].sum()) / args.des_per_step)
        info.update(pump_fraction_batches=fr, batch_des=args.des_per_step)
    sim.close()
    fh = t.facets["nb_mc_hit"][0].astype(float)
    tex_ids = list(getattr(model, "tex_ids", []))
    tex_frac = float(fh[tex_ids].sum() / max(fh.sum(), 1.0)) if tex_ids else 0.0
    b_hit, nodes, prims = algorithmic_bytes_per_hit(model, tex_frac, **ray_geom)
    launches = steps + 1
    launch_s = kernel_ms / 1e3 / launches
    hits_per_launch = hits / launches
    achieved = b_hit * hits_per_launch / launch_s / 1e9
    name = "tube" if tube else "beam_pipe"
    traffic = profile_value(f"trace_{name}_pmc.txt", "traffic_bytes_per_launch")
    out = {
        "metric": "MC hits/s on 1 MI355X (plain tracer, config-1 tube)" if tube else
                  "MC hits/s on 1 MI355X (plain tracer, config-2 beam pipe)",
        "value": value,
        "unit": "hits/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "tube_R1cm_L10cm_64x50_wall_quads" if tube else
                               "beam_pipe_100k_triangles_8_ports_16_textures_4_anglemaps",
                   "facets": model.n_facets, "desorptions_per_step": args.des_per_step,
                   "hits_per_desorption": hits / (steps * args.des_per_step),
                   "seed": 42, "parallelism": "single GPU", "carry": not args.no_carry,
                   "drain_ms": drain_s * 1e3, "launch": tune, "job": "steps x desorptions_per_step, every particle traced to its end"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_hit": b_hit, "bvh_nodes_per_ray": nodes, "facet_tests_per_ray": prims,
                     "kernel_ms_per_launch": launch_s * 1e3,
                     "traffic_source": f"profiles/trace_{name}_pmc.txt" if traffic else
                                       "no summary of this kernel build under profiles/"},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(model, name="tube" if tube else "beam-pipe"),
    }
    # the roof the counters show: VALU issue and memory wait of the trace kernel
    valu = profile_value(f"trace_{name}_sq.txt", "valu_issue_per_cu")
    wait = profile_value(f"trace_{name}_sq.txt", "wait_any_fraction")
    if valu is not None:
        out["issue"] = {"bound": "valu_issue", "achieved": valu, "peak": 1.0,
                        "unit": "wave64 VALU instructions per CU cycle", "frac": valu,
                        "wait_any_fraction": wait, "source": f"profiles/trace_{name}_sq.txt"}
    peak = atomic_peak()
    atom = profile_value(f"trace_{name}_atomics.txt", "atomics_per_launch")
    if peak:
        out["atomics"] = {"bound": "l2_atomics", "achieved": atom / launch_s if atom else None,
                          "peak": peak["f64_per_s"], "unit": "atomic ops/s",
                          "frac": atom / launch_s / peak["f64_per_s"] if atom else None,
                          "atomics_per_hit": atom / hits_per_launch if atom else None,
                          "peak_u64": peak["u64_per_s"], "peak_measured_by": peak["ops"]}
    if tube:
        # facet 1 is the outlet disk: transmission = absorbed there / desorbed
        out["config"]["transmission"] = float(t.facets["nb_abs_equiv"][0][1]) / float(t.glob[0]["nb_desorbed"])
    g = gather_rates()
    if traffic and g:
        # the measured L2-miss traffic against the rate the same access
        # pattern (random 128 B lines) reaches at this working-set size
        rate = traffic / launch_s / 1e9
        out["roofline"]["traffic_GBps"] = rate
        out["roofline"]["random_line_read_GBps"] = g
        out["roofline"]["traffic_frac_of_random_line_rate"] = rate / g["64MiB_GBps"]
    if tube:
        # 3,202 facets: nodes, leaf records, facet data and tallies (about
        # 2 MB) stay in each XCD's 4 MB L2, so the HBM fraction is nominal
        # here; the issue object (VALU issue and memory wait) is the roof
        # that binds
        out["roofline"]["note"] = ("working set L2-resident (~2 MB of nodes, leaf records, facet data and tallies); "
                                   "the HBM fraction is nominal, the issue object is the binding roof")
    if out["cpu_baseline"]:
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    return out, info


def tube_leg(args):
    """The tube job (10 x 1e6 desorptions) and its CPU baseline, as extra
    keys of the default line."""
    a = parse_args(["--workload", "tube"] + (["--no-cpu"] if args.no_cpu else []))
    out, _ = plain_leg(a, True)
    r = {"metric": "MC hits/s, config-1 tube (R 1 cm, L 10 cm, 3,200 wall quads), 1e7 desorptions",
         "value": out["value"], "unit": "hits/s", "kernel_ms_per_launch": out["roofline"]["kernel_ms_per_launch"],
         "transmission": out["config"]["transmission"], "launch": out["config"]["launch"],
         "roofline": out["roofline"], "cpu_baseline": out["cpu_baseline"]}
    if "issue" in out:
        r["issue"] = out["issue"]
    if out["cpu_baseline"]:
        r["speedup_vs_cpu"] = out["speedup_vs_cpu"]
    return r


def matched_sigma(info, cpu):
    """Time for the GPU and for the CPU oracle to bring the standard error
    of the pump-absorbed fraction p to 1e-5. Each particle ends on a pump
    or not, so the per-particle variance is p (1 - p); the GPU's batch
    means give it independently (both are reported)."""
    import numpy as np
    fr = np.array(info["pump_fraction_batches"])
    n = info["batch_des"]
    p = float(fr.mean())
    v_bin = p * (1 - p)
    v_bm = float(fr.var(ddof=1) * n)
    target = 1e-5
    gpu_s = v_bin / target ** 2 / info["des_per_s"]
    cpu_s = v_bin / target ** 2 / cpu["desorptions_per_s"]
    return {"tally": "fraction of desorbed particles absorbed by the 8 pumps", "p": p,
            "sigma_target": target, "gpu_s": gpu_s, "cpu_s": cpu_s, "speedup": cpu_s / gpu_s,
            "sigma_per_sqrt_desorption": {"binomial": math.sqrt(v_bin), "gpu_batch_means": math.sqrt(v_bm),
                                          "batches": len(fr), "desorptions_per_batch": n},
            "cpu_cores": cpu["cores"]}


def progress(msg):
    """One line per finished stage on stderr (long workloads stay visibly alive)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def batched_leg(n_variants, lo, hi, graph=False):
    """Uneven variants of the beam pipe (facet counts log-uniform in 2k-150k,
    NEG sticking log-uniform in 1e-4..1e-1, seeded per variant): one
    persistent launch over all of them (optionally replayed from a captured
    hipGraph) vs one launch per variant, each tracing its desorptions to
    the end. Setup (model upload, BVH build) is outside both timed regions."""
    import numpy as np

    from molflow_sources_amd import geometry
    from molflow_sources_amd.simulation import GpuSimulation
    vs = geometry.variant_models(n_variants, seed=42)
    progress(f"{len(vs)} variants, {sum(m.n_facets for m in vs)} facets")
    rng = np.random.default_rng(7)
    n_des = np.exp(rng.uniform(math.log(lo), math.log(hi), len(vs))).astype(np.uint64)
    seeds = np.array([1_000_003 * (v + 1) + 42 for v in range(len(vs))], dtype=np.uint64)
    max_events = 50_000_000
    sim = GpuSimulation(0)
    sim.load_variants(vs)
    sim.rebuild_accel()
    sim.reset()
    progress("batched accel built")
    tr = sim.autotune(int(n_des.sum() // 8))
    progress(f"autotune: {tr.waves} waves/SIMD, block {tr.block}")
    if graph:
        sim.enable_graph(True)
    sim.reset()
    sim.run_variants(n_des, seeds, max_events=max_events)   # warm-up (and graph capture)
    sim.clear_tallies()
    sim.reset()
    t0 = time.perf_counter()
    r = sim.run_variants(n_des, seeds, max_events=max_events)
    wall_b = time.perf_counter() - t0
    hits_b, kms_b = r.hits, r.kernel_ms
    batched_hits = [int(sim.read(variant=v).glob[0]["nb_mc_hit"]) for v in range(len(vs))]
    progress(f"batched launch: {hits_b} hits in {wall_b:.2f} s")
    sim.close()
    p = GpuSimulation(0)
    p.set_kernel_variant(tr.waves)
    p.set_launch(tr.block, 0)
    wall_p, hits_p, kms_p, same = 0.0, 0, 0.0, True
    for v, m in enumerate(vs):
        p.load(m)
        p.rebuild_accel()
        p.reset(seed=int(seeds[v]))
        t0 = time.perf_counter()
        r = p.run(int(n_des[v]), max_events=max_events)
        wall_p += time.perf_counter() - t0
        hits_p += r.hits
        kms_p += r.kernel_ms
        same &= r.hits == batched_hits[v]
        if v % 32 == 31:
            progress(f"per-variant launches: {v + 1} done")
    p.close()
    out = {
        "metric": "aggregate MC hits/s, uneven beam-pipe variants: one batched launch vs one launch per variant",
        "value": hits_b / wall_b, "unit": "hits/s", "n_gpus": 1, "higher_is_better": True,
        "per_variant_launches_hits_per_s": hits_p / wall_p, "speedup_batched": (hits_b / wall_b) / (hits_p / wall_p),
        "batched_wall_s": wall_b, "per_variant_wall_s": wall_p, "batched_kernel_ms": kms_b,
        "per_variant_kernel_ms": kms_p, "hits": hits_b, "identical_hits_per_variant": bool(same),
        "dtype": "f64", "data": "synthetic", "hipgraph": bool(graph),
        "config": {"workload": "beam_pipe_variants", "variants": len(vs),
                   "facets_total": int(sum(m.n_facets for m in vs)),
                   "facets_min": int(min(m.n_facets for m in vs)), "facets_max": int(max(m.n_facets for m in vs)),
                   "desorptions_total": int(n_des.sum()), "desorptions_per_variant": [lo, hi],
                   "launch": {"waves_per_simd": tr.waves, "block": tr.block, "blocks_per_cu": tr.blocks_per_cu}},
    }
    return out


def adjoint_leg(n, fd_des=0):
    """Sticking adjoint on the beam pipe in low-flux mode: 32 NEG sub-bands
    (base sticking log-uniform in 1e-3..5e-2), 64 target wall facets. Cost
    of a run with the path-score Jacobian vs the same forward run without
    it; with fd_des > 0 also every d(upstream . J0)/ds_g against finite
    differences of the same particles (fd_check)."""
    import numpy as np

    from molflow_sources_amd import geometry
    from molflow_sources_amd.simulation import GpuSimulation
    m = geometry.beam_pipe_model()
    m.settings.low_flux_mode = 1
    m.settings.low_flux_cutoff = 1e-6
    grp, base, tg = geometry.adjoint_setup(m, n_groups=32, n_targets=64, seed=42)
    for g, s in enumerate(base):
        m.facets["sticking"][grp == g] = s
    up = np.random.default_rng(42).normal(size=len(tg))
    max_events = 50_000_000
    a = GpuSimulation(0)
    a.load(m)
    a.rebuild_accel()
    a.adjoint_setup(grp, tg)
    a.adjoint_set_sticking(base)
    ta = a.autotune(max(n // 4, 10_000))
    a.reset(seed=42)
    a.run(min(n, 20_000), max_events=max_events)
    a.reset(seed=42)
    ra = a.run(n, max_events=max_events)
    grad = a.adjoint_backward(up)
    progress(f"adjoint: {ra.hits} hits in {ra.kernel_ms:.1f} ms")
    # forward cost: same model, same sticking, no Jacobian, timed with its
    # own autotuned launch and with the adjoint's; the faster is the baseline
    f = GpuSimulation(0)
    f.load(m)
    f.rebuild_accel()
    tf = f.autotune(max(n // 4, 10_000))
    shapes = [(tf.waves, tf.block, tf.blocks_per_cu, tf.bvh_width, tf.leaf_size),
              (ta.waves, ta.block, ta.blocks_per_cu, ta.bvh_width, ta.leaf_size)]
    fwd = []
    for w, blk, per_cu, width, leaf in shapes:
        f.rebuild_accel(width, leaf)
        f.set_kernel_variant(w)
        f.set_launch(blk, per_cu)
        f.reset(seed=42)
        f.run(min(n, 20_000), max_events=max_events)
        f.reset(seed=42)
        fwd.append(f.run(n, max_events=max_events))
    f.close()
    k = min(range(len(fwd)), key=lambda i: fwd[i].kernel_ms)
    rf = fwd[k]
    tf_used = shapes[k]
    progress(f"forward: {rf.hits} hits in {rf.kernel_ms:.1f} ms (launches tried: "
             + ", ".join(f"{r.kernel_ms:.1f} ms" for r in fwd) + ")")
    out = {
        "metric": "adjoint cost vs one forward run, beam pipe, 32 sticking groups x 64 targets",
        "value": ra.kernel_ms / rf.kernel_ms, "unit": "x forward kernel time", "n_gpus": 1,
        "higher_is_better": False, "forward_kernel_ms": rf.kernel_ms, "adjoint_kernel_ms": ra.kernel_ms,
        "forward_hits_per_s": rf.hits / (rf.kernel_ms / 1e3), "adjoint_hits_per_s": ra.hits / (ra.kernel_ms / 1e3),
        "hits": ra.hits,
        "launch": {"forward": list(tf_used[:2]), "adjoint": [ta.waves, ta.block],
                   "forward_kernel_ms_tried": [r.kernel_ms for r in fwd]},
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "beam_pipe_low_flux_adjoint", "facets": m.n_facets, "desorptions": n,
                   "groups": len(base), "targets": len(tg), "low_flux_cutoff": 1e-6},
    }
    if fd_des > 0:
        out["gradient"] = [float(x) for x in grad]
        out["fd_check"] = fd_check(a, base, up, fd_des, max_events)
    a.close()
    return out


def fd_check(a, base, up, fd_des, max_events, nb=8):
    """Every d(up . J0)/ds_g against finite differences of the same runs
    with common random numbers (the same particles, so the same paths),
    in nb independent particle batches. The differences are Richardson-
    extrapolated central differences at h and 2h (h = 1e-3 s_g), which
    removes the O(h^2) truncation; what remains is the low-flux cutoff's
    boundary term and noise. z divides the mean paired difference by its
    own batch standard error."""
    import numpy as np
    per = max(1, fd_des // nb)

    def objective(s, b):
        a.adjoint_set_sticking(s)
        a.reset(seed=42, particle_offset=b * per)
        a.run(per, max_events=max_events)
        j0, jac = a.adjoint_read()
        return float(up @ j0) / per, (up @ jac) / per

    adj_b = np.array([objective(base, b)[1] for b in range(nb)])     # [batch][group]
    fd_b = np.zeros_like(adj_b)
    for g in range(len(base)):
        h = 1e-3 * base[g]
        for b in range(nb):
            f = {}
            for k in (-2, -1, 1, 2):
                sv = base.copy()
                sv[g] += k * h
                f[k] = objective(sv, b)[0]
            d1 = (f[1] - f[-1]) / (2 * h)
            d2 = (f[2] - f[-2]) / (4 * h)
            fd_b[b, g] = (4 * d1 - d2) / 3
        progress(f"finite differences: group {g + 1}/{len(base)}")
    d = fd_b - adj_b
    sem_d = d.std(axis=0, ddof=1) / math.sqrt(nb)
    z = np.where(sem_d > 0, d.mean(0) / np.where(sem_d > 0, sem_d, 1.0), 0.0)
    rel = np.abs(d.mean(0)) / np.abs(adj_b.mean(0))
    return {"paired_z": [float(x) for x in z], "max_abs_z": float(np.max(np.abs(z))),
            "groups_within_2": int(np.sum(np.abs(z) <= 2)), "groups": len(z),
            "mean_z_squared": float(np.mean(z ** 2)), "max_rel_diff": float(np.max(rel)),
            "difference": "Richardson central differences (h, 2h; h = 1e-3 s_g) minus the path-score "
                          "gradient, same particles", "batches": nb, "desorptions_per_batch": per}


if __name__ == "__main__":
    main()
