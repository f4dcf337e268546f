"""Benchmark of the gfx950 radar integrator on its headline workload.

Workload (one step = one radar frame): a 200,000-triangle vehicle-like clutter mesh (seed 1)
10–30 m ahead, 1 Tx × 4 Rx at λ/2, 77 GHz / 60 MHz/µs / 10 MS/s / 256 samples, maximum depth 4,
path guiding on: 4 training passes of 2^20 paths, then the measured frame of 2^24 paths.
Inputs (mesh BVH, scene, guide) are resident on the device when timing starts.

Prints one JSON line: `value` = the frame's paths (2^24) per second of whole-frame time
(training passes included), beside the paths/s of all passes and of the final pass alone,
frames/s and p50/p99 frame latency; the roofline of the dominant kernel (the closest-hit
traversal of the wavefront pipeline) from hardware counters collected in this run (child
processes of this script under rocprofv3 --pmc, before this process touches the GPU): its
VALU issue rate against the chip's, its HBM bytes per launch (FETCH_SIZE / WRITE_SIZE scaled
by a streaming kernel of known size in the same child run) and its algorithmic BVH node +
triangle bytes (deterministic traversal counters) per HIP-event launch duration; the IF
relative L2 of the frame against the FP64 oracle on a bounded sample with the same trained
guide, hit ids against the oracle's FP64-chain and reference-arithmetic (FP32 chain)
instances; and the oracle's FP32 restatement timed on the host cores (cpu_baseline).
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import hussar_amd as H  # noqa: E402
from hussar_amd import scenes  # noqa: E402
from hussar_amd._lib import (OPT_COUNTERS, OPT_GRAPH, OPT_MARKERS, OPT_PIPELINE, OPT_RECORD_HITS,  # noqa: E402
                             OPT_SHADOW_SPLAT, OPT_TAIL)

METRIC = "paths/s & frames/s at ≤1e-3 IF rel-L2 vs FP64; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0
# four-wide BVH node: 8-bit child boxes for the closest-hit traversal, float boxes for the
# any-hit (shadow) traversal; triangle slot
NODE_BYTES, NODE_BYTES_ANY, TRI_BYTES = 64, 128, 48


def workload(name):
    rf = scenes.dihedral_rf()  # 77 GHz, 60 MHz/us, 10 MS/s, 256 samples (25.6 us)
    if name == "vehicle":
        mesh, _ = scenes.vehicle_mesh(seed=1, triangles=200_000)
        scene = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
        knobs = H.Knobs(maxDepth=4, doGuiding=True)
        fc = H.FrameConfig(1, 256, 4)
        sched = (4, 1 << 20, 1 << 24)
        desc = ("vehicle-like clutter 200k triangles (seed 1) at 10-30 m, 1 Tx x 4 Rx lambda/2, 77 GHz 60 MHz/us "
                "10 MS/s 256 samples, depth 4, guiding 4 x 2^20 training paths + 2^24 frame paths")
    elif name == "dihedral":
        mesh = scenes.dihedral_mesh()
        scene = scenes.dihedral_pose(0.0)
        knobs = H.Knobs(maxDepth=3, doGuiding=False)
        fc = H.FrameConfig(1, 256, 1)
        sched = (0, 0, 1 << 20)
        desc = "dihedral 2 x 10 cm plates at 1.5 m, 1 Tx/Rx, 256 samples, depth 3, 2^20 paths, no guiding"
    else:
        raise SystemExit(f"unknown workload {name}")
    return mesh, scene, knobs, fc, sched, desc


# VALU issue: a CU's 4 SIMDs each issue one wave64 VALU instruction per 4 cycles
SIMDS = 1024  # 256 CUs x 4 SIMDs
# the byte-counter calibration kernel: the frame normalisation of a 64 x 4096 x 16 frame reads
# its sum and compensation (2 x 2^23 doubles) and writes 2^23 doubles
CALIB_FRAME = (64, 4096, 16)
CALIB_READ, CALIB_WRITE = 2 * 8 * (1 << 23), 8 * (1 << 23)
PMC_PASSES = [("fetch", ["FETCH_SIZE"]), ("write", ["WRITE_SIZE"]),
              ("valu", ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "GRBM_GUI_ACTIVE"]),
              ("texture", ["TA_TA_BUSY_sum", "TD_TD_BUSY_sum", "GRBM_GUI_ACTIVE"])]
CUS = 256
# kernel name patterns per key (a dispatch counts when its name contains one of them)
PMC_KERNELS = {"trav": ("k_wf_trav<false>",), "shadow": ("k_wf_shadow<false>", "k_wf_shadow_splat<false,"),
               "calib": ("k_post_normalize",)}


def counter_child(args):
    """Run under rocprofv3 by collect_counters: the calibration kernel of known size, then one
    step of the workload (no counting step)."""
    cal = H.HipBackend(scenes.dihedral_mesh())
    cal._configure(H.FrameConfig(*CALIB_FRAME))
    cal.peak(False)
    del cal
    mesh, scene, knobs, fc, sched, _ = workload(args.workload)
    pt = H.PathTracer(**{k: getattr(knobs, k) for k in knobs.__dataclass_fields__})
    pt.configureFrame(fc)
    backend = H.HipBackend(mesh, pt)
    backend.set_option(OPT_PIPELINE, args.pipeline)
    backend.set_option(OPT_SHADOW_SPLAT, args.shadow_splat)
    if args.tail >= 0:
        backend.set_option(OPT_TAIL, args.tail)
    run_step(backend, pt, scene, sched)
    backend.fetch()


def collect_counters(args):
    """Sums of each counter over the dispatches of the kernels in PMC_KERNELS, from child runs
    of this script under rocprofv3 --pmc, one run per counter group (rocprofv3 does not split
    counters over passes). Returns (counters, None) or (None, reason)."""
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    for name, ctrs in PMC_PASSES:
        d = tempfile.mkdtemp(prefix="hsr_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "240", exe, "--pmc", *ctrs, "--output-format", "csv", "-d", d, "-o", "pmc",
               "--", sys.executable, os.path.abspath(__file__), "--counter-child", "--workload", args.workload,
               "--pipeline", str(args.pipeline), "--shadow-splat", str(args.shadow_splat), "--tail", str(args.tail)]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True):
            with open(f, newline="") as fh:
                rows += list(csv.DictReader(fh))
        shutil.rmtree(d, ignore_errors=True)
        print(f"counters: pass {name} ({' '.join(ctrs)}): rc {r.returncode}, {len(rows)} rows, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if not rows:
            return None, f"pass {name}: rc {r.returncode}, no counter rows: {(r.stderr or '')[-300:]}"
        for key, pats in PMC_KERNELS.items():
            sel = [q for q in rows if any(pat in q.get("Kernel_Name", "") for pat in pats)]
            if not sel:
                return None, f"pass {name}: no dispatch of {' or '.join(pats)}"
            ent = res.setdefault(key, {})
            ent["dispatches"] = len({q.get("Dispatch_Id") for q in sel})
            for c in ctrs:
                # each pass's own busy cycles under their own key (they come from another run)
                k = c if c != "GRBM_GUI_ACTIVE" or name == "valu" else f"{c}:{name}"
                ent[k] = sum(float(q["Counter_Value"]) for q in sel if q.get("Counter_Name") == c)
    return res, None


def roofline(pipeline, acc, cnt, steps, pmc=None, pmc_error=None, clock_mhz=2400):
    """Roofline of the dominant kernel, the closest-hit traversal (one launch per bounce), per
    launch: its HIP-event duration on the backend's stream averaged over the timed steps, and
     * algorithmic: BVH node + triangle bytes the traversal must read (deterministic counters of
       one step) per duration, against HBM;
     * hbm: HBM bytes per launch measured by FETCH_SIZE and WRITE_SIZE, scaled by what they read
       for the calibration kernel of known size, per duration, against HBM;
     * valu: VALU wave instructions per launch (SQ_INSTS_VALU) per duration, against the SIMDs'
       issue rate (SIMDS x clock / 4);
     * texture: the busy shares of the CUs' texture address and data units (TA_TA_BUSY,
       TD_TD_BUSY), which every vector memory load passes.
    `bound` names the largest measured fraction; the top-level achieved / peak / frac are its."""
    if pipeline == 1:
        # the closest-hit traversal launches' own visits: not the shadow traversal's nor those of
        # the fused tail (the last bounces, one launch per batch)
        n = cnt["nodes_visited"] - cnt["nodes_shadow"] - cnt["nodes_tail"]
        t = cnt["tris_tested"] - cnt["tris_shadow"] - cnt["tris_tail"]
        launches = cnt["closest_launches"]
        kms = acc["closest_ms"] / acc["closest_launches"]
        kernel, sym = "k_wf_trav<false> (closest-hit traversal, one launch per bounce)", "k_wf_trav_closest"
        rays = cnt["rays_closest"]
        sh_bytes = (cnt["nodes_shadow"] * NODE_BYTES_ANY + cnt["tris_shadow"] * TRI_BYTES) / cnt["shadow_launches"]
        sh_ms = acc["shadow_ms"] / acc["shadow_launches"]
        second = {"kernel": "k_wf_shadow<false> and k_wf_shadow_splat<false, *> (packet any-hit traversal; the "
                            "second also splats, in batches without guiding records)", "kernel_ms": sh_ms,
                  "algorithmic_bytes": sh_bytes, "achieved": sh_bytes / (sh_ms * 1e-3) / 1e9,
                  "rays_shadow": cnt["rays_shadow"], "nodes_visited": cnt["nodes_shadow"],
                  "tris_tested": cnt["tris_shadow"]}
    else:
        n, t = cnt["nodes_visited"], cnt["tris_tested"]
        launches = cnt["launches"]
        kms = acc["trace_ms"] / acc["launches"]
        kernel, sym = "k_trace<false> (fused path kernel)", "k_trace"
        rays = cnt["rays_closest"] + cnt["rays_shadow"]
        second = None
    alg = (n * NODE_BYTES + t * TRI_BYTES) / launches
    alg_gbs = alg / (kms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": alg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": alg_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": kernel, "kernel_symbol": sym,
           "kernel_ms": kms, "launches_per_step": launches,
           "algorithmic": {"bytes_per_launch": alg, "achieved": alg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": alg_gbs / HBM_PEAK_GBS,
                           "note": "node visits x 64 B + triangle tests x 48 B; the BVH is served by the L1/L2"},
           "nodes_visited": n, "tris_tested": t, "rays": rays, "events": cnt["events"]}
    if pipeline == 1 and pmc:
        tr, cal = pmc["trav"], pmc["calib"]
        f_read = CALIB_READ / (cal["FETCH_SIZE"] * 1024.0 / cal["dispatches"])
        f_write = CALIB_WRITE / (cal["WRITE_SIZE"] * 1024.0 / cal["dispatches"])
        traffic = (tr["FETCH_SIZE"] * f_read + tr["WRITE_SIZE"] * f_write) * 1024.0 / tr["dispatches"]
        hbm_gbs = traffic / (kms * 1e-3) / 1e9
        insts = tr["SQ_INSTS_VALU"] / tr["dispatches"]
        valu_rate = insts / (kms * 1e-3) / 1e9
        valu_peak = SIMDS * clock_mhz * 1e6 / 4 / 1e9
        lane = tr["SQ_THREAD_CYCLES_VALU"] / (64.0 * tr["SQ_ACTIVE_INST_VALU"])
        busy = tr["SQ_ACTIVE_INST_VALU"] * 4.0 / SIMDS / (tr["GRBM_GUI_ACTIVE"] / 8.0)
        visits = (n + t) / launches
        sh = pmc["shadow"]
        out["traffic"] = traffic
        out["hbm"] = {"bytes_per_launch": traffic, "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": hbm_gbs / HBM_PEAK_GBS,
                      "calibration": f"FETCH_SIZE x {f_read:.4f} KiB + WRITE_SIZE x {f_write:.4f} KiB (frame "
                                     f"normalisation of {CALIB_FRAME}: {CALIB_READ >> 20} MiB read, "
                                     f"{CALIB_WRITE >> 20} MiB written)"}
        out["valu"] = {"instructions_per_launch": insts, "achieved": valu_rate, "peak": valu_peak,
                       "unit": "G wave64 VALU instructions/s", "frac": valu_rate / valu_peak,
                       "issue_busy": busy, "lane_activity": lane,
                       "instructions_per_visit": insts * 64.0 * lane / visits,
                       "note": "instructions per visit = active-lane VALU instructions per node visit or "
                               "triangle test"}
        cyc = tr["GRBM_GUI_ACTIVE:texture"] / 8.0  # the kernel's busy cycles (the counter sums 8 XCDs)
        out["texture"] = {"ta_busy": tr["TA_TA_BUSY_sum"] / CUS / cyc, "td_busy": tr["TD_TD_BUSY_sum"] / CUS / cyc,
                          "note": "share of the kernel's cycles in which a CU's texture address (TA) and texture "
                                  "data (TD) units are busy: every node and triangle load passes both, so a "
                                  "high TD share beside a high VALU share means both bind"}
        out["shadow_kernel_counters"] = {
            "hbm_bytes_per_launch": (sh["FETCH_SIZE"] * f_read + sh["WRITE_SIZE"] * f_write) * 1024.0 /
            sh["dispatches"],
            "valu_instructions_per_launch": sh["SQ_INSTS_VALU"] / sh["dispatches"],
            "lane_activity": sh["SQ_THREAD_CYCLES_VALU"] / (64.0 * sh["SQ_ACTIVE_INST_VALU"])}
        best = max(("valu", out["valu"]), ("hbm", out["hbm"]), key=lambda kv: kv[1]["frac"])
        out["bound"] = best[0]
        out["achieved"], out["peak"], out["unit"], out["frac"] = (best[1]["achieved"], best[1]["peak"],
                                                                  best[1]["unit"], best[1]["frac"])
    elif pmc_error:
        out["counters_error"] = pmc_error
    if second:
        out["shadow_kernel"] = second
    return out


def run_step(backend, pt, scene, sched):
    passes, train, final = sched
    pt.run_schedule(backend, scene, passes, train, final)


def cpu_baseline(mesh, scene, knobs, fc, sched, seconds, threads):
    """The oracle's FP32 restatement (reference arithmetic, own BVH, std::thread pool with
    256-index batches) on a bounded sample of the same schedule, scaled down by 2^k."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as ob
    om = ob.OracleMesh.from_mesh(mesh)
    passes, train, final = sched

    def timed(scale):
        tr = max(train // scale, 1024) if passes else 0
        t0 = time.perf_counter()
        ob.path_tracer_run_schedule(om, scene, knobs, fc, passes, tr, max(final // scale, 1024),
                                    prec=ob.PREC_REF32, guide=ob.OracleGuide() if knobs.doGuiding else None,
                                    threads=threads)
        return time.perf_counter() - t0, passes * tr + max(final // scale, 1024)

    scale = 256
    dt, paths = timed(scale)
    if dt < seconds / 2:  # grow the sample to about `seconds` of CPU work
        scale = max(1, int(2 ** np.floor(np.log2(scale * dt / seconds))))
        dt, paths = timed(scale)
    return {"value": paths / dt, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"same scene and schedule with path counts / {scale} ({paths} paths, {dt:.1f} s), "
                      "oracle FP32 restatement (reference arithmetic, own SAH BVH2, std::thread pool, "
                      "256-index batches)"}


def host_threads():
    """Threads of the CPU baseline: every core this process may run on, capped by the CPU share
    the host gives the process when it states one (OMP_NUM_THREADS; a GPU box allots 16 cores per
    GPU). Returns (threads, host description)."""
    aff = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(aff, share) if share > 0 else aff
    return threads, {"os_cpu_count": os.cpu_count(), "affinity_cores": aff,
                     "cpu_share": share if share > 0 else None}


def accuracy_check(backend, pt, mesh, scene, knobs, fc, n, threads):
    """Frame of n final-pass paths with the trained sampling tree, GPU vs FP64 oracle (same
    tree, same Halton indices), plus bit-exact hit ids against the same-precision oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as ob
    om = ob.OracleMesh.from_mesh(mesh)
    og = None
    guide = pt.guiding if knobs.doGuiding else None
    if guide is not None:
        child, p0 = guide.sampling()
        og = ob.OracleGuide()
        og.set_sampling(child, p0)
    backend.clear()
    rec = n
    backend.set_option(OPT_RECORD_HITS, rec)
    backend.run_raw(scene, knobs, 0, n, guide=guide, is_final=True)
    backend.set_option(OPT_RECORD_HITS, 0)
    z, tw = backend.fetch("complex128")
    hits, nee = backend.hit_record(rec, knobs.maxDepth, len(scene.receivers()))
    zf, twf, _, _ = ob.render(om, scene, knobs, fc, prec=ob.PREC_FP64, count=n, is_final=True, guide=og,
                              threads=threads)
    zm, twm, recm, _ = ob.render(om, scene, knobs, fc, prec=ob.PREC_MIXED, count=rec, is_final=True, guide=og,
                                 threads=threads, record_paths=rec)
    _, _, rec32, _ = ob.render(om, scene, knobs, fc, prec=ob.PREC_REF32, count=rec, is_final=True, guide=og,
                               threads=threads, record_paths=rec)
    zf = zf / twf
    rl2 = float(np.linalg.norm(z - zf) / np.linalg.norm(zf))
    peaks_gpu = np.abs(z).reshape(-1, fc.channelCount).argmax(0).tolist()
    peaks_ref = np.abs(zf).reshape(-1, fc.channelCount).argmax(0).tolist()
    hit_paths = int((hits[:, 0] >= 0).sum())
    return {"if_rel_l2_vs_fp64": rl2, "paths": n, "peak_bins_equal": peaks_gpu == peaks_ref,
            "hit_ids_bitexact": bool(np.array_equal(hits, recm[0]) and np.array_equal(nee, recm[1])),
            "hit_paths_compared": rec, "paths_with_a_hit": hit_paths,
            "vs_reference_arithmetic": {
                "note": "oracle instance with the reference's FP32 geometry chain (ref32); the product "
                        "traces rays of an FP64 chain rounded to float",
                "paths_with_other_hit_ids": int((hits != rec32[0]).any(1).sum()),
                "paths_with_other_nee": int((nee != rec32[1]).reshape(rec, -1).any(1).sum())}}


def serving_run(srv, ids, trace, rf, knobs, rate, guide_mode):
    """The trace replayed open-loop at `rate` requests/s (each request submitted at its arrival
    time); returns the server's statistics over it."""
    reqs = []
    for q in trace:
        sc = scenes.radar_array(rf, channels=q["channels"], position=q["pos"])
        reqs.append((ids[q["scene"]], sc, H.FrameConfig(1, q["bins"], q["channels"]), q["paths"]))
    srv.reset_stats()
    t0 = time.perf_counter()
    tickets = []
    for q, (sid, sc, fc, n) in zip(trace, reqs):
        delay = t0 + q["t"] * 100.0 / rate - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(sid, sc, knobs, fc, n, guide_mode=guide_mode))
    for t in tickets:
        srv.wait(t, fetch=False)
    st = srv.stats()
    return {"frames_per_s": st["frames_per_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
            "mean_ms": st["mean_ms"], "completed": st["completed"], "rejected": st["rejected"],
            "memory_waits": st["memory_waits"]}


def serving_bench(args):
    """The serving mix replayed open-loop at 50, 100, 200 and 500 requests/s in three forms:
     * plain: one frame at a time (1 slot), each request the reference's PathTracer::run with
       its doubling guiding schedule (training iterations, then the final iteration);
     * concurrent: `--slots` slots, still PathTracer::run per request;
     * served: `--slots` slots, each scene's resident guiding tree (trained once per scene,
       before the trace, by one request at the scene's nominal pose) and every request's whole
       budget as one final pass with it;
     * served_graph: the same with every batch replayed from its launch shape's graph
       (HSR_OPT_GRAPH) instead of launched stage by stage as in the other forms.
    Frames/s, p50/p99 latency, rejections and memory waits of each."""
    from hussar_amd.serving import GUIDE_PER_REQUEST, GUIDE_SCENE, GUIDE_TRAIN, Server, request_mix
    rf = scenes.dihedral_rf()
    meshes = [scenes.vehicle_mesh(seed=10 + k, triangles=n)[0] for k, n in enumerate((10_000, 300_000, 2_000_000))]
    trace = request_mix(args.requests, 100.0)  # arrival times at 100/s, rescaled per rate
    knobs = H.Knobs(maxDepth=4, doGuiding=True)
    rates = [args.rate] if args.rate > 0 else [50.0, 100.0, 200.0, 500.0]
    forms = {"plain": (1, GUIDE_PER_REQUEST, False), "concurrent": (args.slots, GUIDE_PER_REQUEST, False),
             "served": (args.slots, GUIDE_SCENE, False), "served_graph": (args.slots, GUIDE_SCENE, True)}
    results = {name: {} for name in forms}
    for name, (slots, mode, graphs) in forms.items():
        srv = Server(slots=slots, graphs=graphs)
        ids = [srv.open_scene(m) for m in meshes]
        nominal = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
        for sid in ids:  # warm-up, and the scenes' resident trees
            srv.wait(srv.submit(sid, nominal, knobs, H.FrameConfig(1, 256, 4), 1 << 20,
                                guide_mode=GUIDE_TRAIN if mode == GUIDE_SCENE else GUIDE_PER_REQUEST), fetch=False)
        for rate in rates:
            results[name][str(int(rate))] = serving_run(srv, ids, trace, rf, knobs, rate, mode)
            print(f"serving: {name} at {rate:.0f}/s: {results[name][str(int(rate))]['frames_per_s']:.1f} frames/s",
                  file=sys.stderr, flush=True)
        srv.close()
    paths = sum(q["paths"] for q in trace)
    key = str(int(rates[-1]))
    served, plain = results["served"][key], results["plain"][key]
    return {
        "metric": "serving frames/s on the mixed-request trace (one GPU)", "value": served["frames_per_s"],
        "unit": "frames/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "data": "synthetic",
        "dtype": "f32 traversal+fields / f64 geometry chain+phase+sums",
        "config": {"workload": "serving mix: scenes of 10k / 300k / 2M triangles, Poisson arrivals (seed 3) "
                               "at each rate, pose within 20 m, bins U{128..512}, receivers U{1..12}, paths "
                               "2^U{18..24}, depth 4, guiding on",
                   "requests": args.requests, "paths_total": paths, "slots": args.slots, "rates": rates},
        "latency_ms": {"p50": served["p50_ms"], "p99": served["p99_ms"], "mean": served["mean_ms"]},
        "speedup_vs_plain": {r: results["served"][r]["frames_per_s"] / results["plain"][r]["frames_per_s"]
                             for r in results["served"]},
        "forms": results,
    }


def target_bins(points, rf, N):
    """Range bins N·frac((L/c)·S/f_s) of single-bounce returns from points seen by a radar at
    the origin (round trip L = 2|p|, radar.h:232-241), rounded half away from zero, modulo N."""
    c = float(np.float32(299792458.0))
    out = []
    for p in points:
        x = N * (((2.0 * np.linalg.norm(p) / c) * rf.freqSlope / rf.adcRate) % 1.0)
        out.append(int(np.floor(x + 0.5)) % N)
    return out


def precision_bench(args):
    """Targets at 5, 50 and 150 m over a 1M-triangle rough ground, 1 Tx x 4 Rx, 512 bins,
    2^24 paths, depth 4: throughput of each IF-phase variant, and against the FP64 oracle on a
    bounded sample (the first 2^16 paths, the same Halton indices): its relative L2, the peak
    range bin of every channel, and the phase error at the bins of the six targets (a plate and
    a trihedral per range; at 10 MS/s the 50 and 150 m targets alias into bins near 0)."""
    from hussar_amd._lib import OPT_PHASE_MODE
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as ob
    mesh = scenes.precision_mesh(seed=2)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.0, 0.0))
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    fc = H.FrameConfig(1, 512, 4)
    n, sample = 1 << 24, 1 << 16
    b = H.HipBackend(mesh)
    b._configure(fc)
    om = ob.OracleMesh.from_mesh(mesh)
    zf, twf, _, _ = ob.render(om, sc, kn, fc, prec=ob.PREC_FP64, count=sample, threads=host_threads()[0])
    zf = (zf / twf).reshape(512, 4)
    pts = []
    for rng in (5.0, 50.0, 150.0):  # plate centre and trihedral corner (scenes.precision_mesh)
        pts += [np.array([-1.5, 0.0, rng]), np.array([1.5, -0.25, rng + 0.25])]
    bins = sorted(set(target_bins(pts, rf, 512)))
    out = {}
    for mode, name in ((0, "fp64"), (2, "double-float"), (1, "fp32")):
        b.set_option(OPT_PHASE_MODE, mode)
        b.clear()
        b.run_raw(sc, kn, 0, sample)
        z, _ = b.fetch("complex128")
        z = z.reshape(512, 4)
        rl2 = float(np.linalg.norm(z - zf) / np.linalg.norm(zf))
        peaks, peaks_ref = np.abs(z).argmax(0).tolist(), np.abs(zf).argmax(0).tolist()
        dphi = np.angle(z[bins] / zf[bins])
        b.clear()
        b.run_raw(sc, kn, 0, n)  # warm-up
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.clear()
            b.run_raw(sc, kn, 0, n)
        b.fetch()
        dt = (time.perf_counter() - t0) / args.steps
        out[name] = {"paths_per_s": n / dt, "ms_per_frame": 1e3 * dt, "if_rel_l2_vs_fp64": rl2,
                     "peak_bins": peaks, "peak_bins_fp64": peaks_ref,
                     "peak_bin_error": [abs(a - c) for a, c in zip(peaks, peaks_ref)],
                     "phase_error_rad_max": float(np.abs(dphi).max()),
                     "phase_error_rad_per_bin": {int(bb): float(np.abs(dphi[k]).max()) for k, bb in enumerate(bins)}}
    return {"metric": "paths/s & IF rel-L2 vs FP64 per phase variant", "value": out["fp64"]["paths_per_s"],
            "unit": "paths/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "data": "synthetic",
            "config": {"workload": "plates + trihedrals at 5/50/150 m over a 1M-triangle rough ground (seed 2), "
                                   "1 Tx x 4 Rx, 512 bins, 2^24 paths, depth 4, no guiding",
                       "triangles": int(mesh.triangle_count()), "accuracy_sample_paths": sample,
                       "target_bins": bins},
            "variants": out}


def urban_bench(args):
    """Urban MIMO: a 2M-triangle street canyon with four moving vehicle blocks, 3 Tx x 4 Rx
    (12 virtual channels), 64 chirps x 256 samples, depth 5, guiding on (4 training passes of
    2^20 paths, then the frame of --paths paths); the frame is the range-Doppler cube with every
    contribution at its own Doppler (the separable 33 x 33 kernel). Reports the frame's paths
    per second of whole-frame time, and the cube's relative L2 against the FP64 oracle on a
    bounded sample (2^13 final-pass paths with the trained tree shared by both sides)."""
    from hussar_amd._lib import OPT_COUNTERS
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bindings as ob
    mesh, vel = scenes.urban_mesh(seed=4)
    rf = scenes.dihedral_rf()
    sc = scenes.mimo_array(rf, tx=3, rx=4, position=(0.0, 1.0, 0.0))
    kn = H.Knobs(maxDepth=5, doGuiding=True)
    fc = H.FrameConfig(64, 256, 12)
    n = args.paths
    sched = (4, 1 << 20, n)
    pt = H.PathTracer(**{k: getattr(kn, k) for k in kn.__dataclass_fields__})
    pt.configureFrame(fc)
    b = H.HipBackend(mesh, pt)
    b.set_velocities(vel)
    b.set_option(OPT_SHADOW_SPLAT, args.shadow_splat)
    if args.tail >= 0:
        b.set_option(OPT_TAIL, args.tail)
    run_step(b, pt, sc, sched)  # warm-up, and the trained tree of the accuracy sample
    child, p0 = pt.guiding.sampling()
    sample = 1 << 13
    b.clear()
    b.run_raw(sc, kn, 0, sample, guide=pt.guiding, is_final=True)
    z, _ = b.fetch("complex128")
    om = ob.OracleMesh.from_mesh(mesh)
    om.set_velocities(vel)
    og = ob.OracleGuide()
    og.set_sampling(child, p0)
    zf, twf, _, _ = ob.render(om, sc, kn, fc, prec=ob.PREC_FP64, count=sample, is_final=True, guide=og,
                              threads=host_threads()[0])
    zf = zf / twf
    rl2 = float(np.linalg.norm(z.reshape(zf.shape) - zf) / np.linalg.norm(zf))
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        run_step(b, pt, sc, sched)
        lat.append(time.perf_counter() - t)
    b.fetch()
    dt = (time.perf_counter() - t0) / args.steps
    st = b.stats()
    b.set_option(OPT_COUNTERS, 1)
    b.clear()
    b.run_raw(sc, kn, 0, 1 << 22, guide=pt.guiding, is_final=True)
    cnt = b.stats()
    return {"metric": "paths/s (range-Doppler cube, MIMO)", "value": n / dt, "unit": "paths/s", "n_gpus": 1,
            "higher_is_better": True, "vs_baseline": None, "data": "synthetic", "ms_per_frame": 1e3 * dt,
            "paths_all_passes_per_s": (4 * (1 << 20) + n) / dt,
            "frame_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                 "p99": 1e3 * float(np.percentile(lat, 99))},
            "config": {"workload": "urban street canyon, 2M triangles (seed 4) with 4 moving vehicle blocks "
                                   "(U[0,20] m/s), 3 Tx x 4 Rx, 64 chirps x 256 samples x 12 channels, depth 5, "
                                   "guiding 4 x 2^20 training paths + the frame",
                       "triangles": int(mesh.triangle_count()), "paths_per_frame": n, "transmitters": 3},
            "trace_ms": st["trace_ms"], "closest_ms": st["closest_ms"], "shadow_ms": st["shadow_ms"],
            "events_per_path": cnt["events"] / float(1 << 22),
            "accuracy": {"cube_rel_l2_vs_fp64": rl2, "paths": sample}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="vehicle")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--pipeline", type=int, default=1, help="0 fused path kernel, 1 wavefront stages")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: request arrivals per second (0: 50, 100, 200 and 500 in turn)")
    ap.add_argument("--paths", type=int, default=1 << 26, help="urban: paths per frame")
    ap.add_argument("--requests", type=int, default=96, help="serving: requests in the trace")
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent slots")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--tail", type=int, default=-1,
                    help="first depth of the fused tail (HSR_OPT_TAIL), 0 none, -1 the library's default")
    ap.add_argument("--graph", type=int, default=0, choices=(0, 1),
                    help="1 replay each batch shape's graph (HSR_OPT_GRAPH), 0 stage-by-stage launches")
    ap.add_argument("--shadow-splat", type=int, default=1, choices=(0, 1),
                    help="1 the shadow stages of batches without guiding records splat (HSR_OPT_SHADOW_SPLAT), "
                         "0 a gather stage after each")
    ap.add_argument("--markers", action="store_true",
                    help="ROCTx ranges around passes, batches and stage launches (HSR_OPT_MARKERS), for "
                         "rocprofv3 --marker-trace")
    ap.add_argument("--counter-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.counter_child:
        counter_child(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the integrator has no multi-GPU path in this build "
              "(paths shard by Halton index; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    threads, host = host_threads()
    if args.workload == "serving":
        print(json.dumps(serving_bench(args)), flush=True)
        return
    if args.workload == "precision":
        print(json.dumps(precision_bench(args)), flush=True)
        return
    if args.workload == "urban":
        print(json.dumps(urban_bench(args)), flush=True)
        return

    # hardware counters first, in child processes, before this process touches the GPU
    pmc, pmc_err = None, "skipped (--no-counters)"
    if not args.no_counters and args.pipeline == 1:
        pmc, pmc_err = collect_counters(args)
    mesh, scene, knobs, fc, sched, desc = workload(args.workload)
    t0 = time.perf_counter()
    pt = H.PathTracer(**{k: getattr(knobs, k) for k in knobs.__dataclass_fields__})
    pt.configureFrame(fc)
    backend = H.HipBackend(mesh, pt)
    backend.set_option(OPT_PIPELINE, args.pipeline)
    backend.set_option(OPT_GRAPH, args.graph)
    backend.set_option(OPT_SHADOW_SPLAT, args.shadow_splat)
    if args.tail >= 0:
        backend.set_option(OPT_TAIL, args.tail)
    if args.markers:
        backend.set_option(OPT_MARKERS, 1)
    bvh_ms = backend.stats()["bvh_build_ms"]
    setup_s = time.perf_counter() - t0

    for _ in range(args.warmup):
        run_step(backend, pt, scene, sched)
    lat = []
    acc = {"trace_ms": 0.0, "launches": 0, "closest_ms": 0.0, "closest_launches": 0, "shadow_ms": 0.0,
           "shadow_launches": 0, "graph_launches": 0}
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        run_step(backend, pt, scene, sched)
        lat.append(time.perf_counter() - t)
        st = backend.stats()  # HIP events on the backend's stream, summed over the step's passes
        for k in acc:
            acc[k] += st[k]
    elapsed = time.perf_counter() - t_start
    passes, train, final = sched
    paths_per_step = passes * train + final
    # the frame's paths per second of whole-frame time (training passes included)
    value = final * args.steps / elapsed
    all_passes = paths_per_step * args.steps / elapsed

    # deterministic traversal counters of one step (separate, untimed instance)
    backend.set_option(OPT_COUNTERS, 1)
    run_step(backend, pt, scene, sched)
    cnt = backend.stats()
    backend.set_option(OPT_COUNTERS, 0)
    roof = roofline(args.pipeline, acc, cnt, args.steps, pmc, pmc_err, H.device_info(0)["clock_mhz"])

    out = {
        "metric": METRIC, "value": value, "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 traversal+fields / f64 geometry chain+phase+sums",
        "data": "synthetic (seeded procedural mesh)",
        "config": {"workload": desc, "triangles": int(mesh.triangle_count()), "paths_per_frame": paths_per_step,
                   "frame_paths": final, "training": f"{passes} x {train}", "max_depth": knobs.maxDepth,
                   "channels": fc.channelCount, "samples": fc.samplesPerChirp, "parallelism": "1 GPU"},
        "paths_all_passes_per_s": all_passes,
        "frames_per_s": args.steps / elapsed,
        "frame_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "pipeline": ["fused path kernel", "wavefront"][args.pipeline],
        "batch_launch": f"graph replay ({acc['graph_launches']} of {acc['launches']} timed batches)" if args.graph
        else "stage by stage",
        "kernel_paths_per_s": paths_per_step * args.steps / (acc["trace_ms"] * 1e-3),
        "host": host,
        "roofline": roof,
        "bvh_build_ms": bvh_ms, "setup_s": setup_s,
    }
    if not args.no_check:
        out["accuracy"] = accuracy_check(backend, pt, mesh, scene, knobs, fc, 1 << 18, threads)
    if not args.no_cpu:
        cb = cpu_baseline(mesh, scene, knobs, fc, sched, args.cpu_seconds, threads)
        # the same metric as value: frame paths per second of whole-frame time
        cb["paths_all_passes_per_s"] = cb["value"]
        cb["value"] = cb["value"] * final / paths_per_step
        out["cpu_baseline"] = cb
        out["gpu_vs_cpu"] = value / cb["value"]
        # a bound for a host whose every core ran the restatement with perfect scaling
        out["gpu_vs_cpu_all_host_cores_linear"] = value / (cb["value"] * (os.cpu_count() or threads) / threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
