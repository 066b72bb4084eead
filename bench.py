#!/usr/bin/env python3
"""Throughput of the MPM snow substep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-autotune] [--no-secondary]
                    [--workload slab4m|slab4m_landing|batch64|adjoint16]

Workload: the 4,194,304-particle snow slab on a 256^3 grid (dx = 1/256,
4x4x4 node blocks), per-particle E0 in [1e5, 2e5], dt = 2.5e-5 — the
configuration the headline metric is quoted on.  A step is one substep over
all particles (binning, stress + F update + P2G, grid update + boundary
conditions, G2P + advection), inputs resident in HBM.  Prints one JSON line.

value: K substeps as the library runs them (HIP-graph segments of one re-bin
+ resort_every substeps), wall clock between stream synchronisations.  The
untimed warmup ends on a re-bin boundary, so the timed window starts with a
re-bin and holds ceil(K / resort_every) of them: binning is always charged,
at least at its amortised rate.  Before timing, the library's autotuner picks
the re-bin interval and block-pass width (reported under "autotune"; results
do not depend on either); --no-autotune times the defaults (16 substeps,
2,048 workgroups).
roofline: the dominant kernel's algorithmic bytes per launch (per-particle
bytes x particles, see alg_bytes_per_particle) over its mean launch time,
measured with HIP events on the library's stream around every launch of a
second pass of K substeps launched directly; traffic and the VALU issue
ceiling come from the committed rocprofv3 counter summaries.
deforming: the same measurement on the slab landing on the sticky floor
(F_E off the identity, the SVD rotating, the clamp active).
cpu_baseline: the oracle (a CPU restatement of the reference path) timed on
this host on the same slab: serial like the reference, and on all the host
cores this process may use, with the calibration to the reference's own
speed per core.
batch64, adjoint16: the secondary workloads, measured in the same run as
their own --workload runs measure them (64 scenes batched against one at a
time, 500 substeps; forward + reverse mode, 200 substeps); --no-secondary
leaves them out.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "realtime-deformations_amd", "python"))

HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (one wave64 vector
# instruction per 4 clocks per SIMD; FP32 unpacked and FP64 alike)
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9


def alg_bytes_per_particle(kernel: str, a_over_n: float) -> float:
    """Algorithmic HBM bytes per particle of one launch of `kernel`.

    Particle state: x, v (12 B each), B, F_E, F_P (36 B each), constants
    m, V0, E0 (4 B each); grid: m + p (16 B) written by P2G and read by the
    grid update, v (12 B) written by the grid update and read by G2P, per
    active node (A of them for N particles).
    """
    if kernel == "fused":
        # one particle-substep: x, v, B, F_E, F_P read and written once each,
        # m, V0, E0 read (v and B stay in registers between the substeps it
        # fuses, but the metric counts them), plus P2G's m+p writes and
        # G2P's v reads on the active nodes
        return 2 * (12 + 12 + 36 + 36 + 36) + 12 + (16 + 12) * a_over_n
    if kernel == "p2g":  # first substep: read x v B F_E F_P m V0 E0, write F_E F_P, write m+p
        return (12 + 12 + 36 * 3 + 12) + 72 + 16 * a_over_n
    if kernel == "g2p":  # last substep: read x and grid v, write x v B
        return 12 + 12 * a_over_n + (12 + 12 + 36)
    if kernel == "grid":  # read m+p, write v
        return (16 + 12) * a_over_n
    if kernel == "bin":  # keys, 3 radix passes of (key, index), 36 fields + caller index moved
        return 12 + 8 + 3 * (4 + 8 + 8) + 8 + 2 * (NFIELDS_BYTES + 4)
    raise KeyError(kernel)


NFIELDS_BYTES = 36 * 4


def substep_alg_bytes(a_over_n: float) -> float:
    """B_alg = 2(12+12+36+36+36) + 4*3 + 56 A/N per particle-substep."""
    return 2 * (12 + 12 + 36 + 36 + 36) + 4 * 3 + 56 * a_over_n


def _note(msg):
    """A progress line on stderr after each stage (stdout holds only the
    JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _host_threads():
    """Host cores this process may use, capped at 16 (one GPU's share of the
    host)."""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline(sc, V0, threads):
    """The oracle (the CPU restatement of the reference path, FP32) on the
    same slab from the same state: two substeps on one thread, as the
    reference runs, then four on `threads` OpenMP threads; about 25 s of host
    time in all.  oracle/calibration.json converts the rate per core into the
    reference's own speed (the oracle against the compiled reference's 9.8e4
    particle-substeps/s per core, measured on the same workload)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    prm = sc["params"]
    n = sc["x"].shape[0]

    def run(t, steps):
        p = O.make_params(sc["dims"], sc["h"], compose=O.LITERAL, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"],
                          xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"], threads=t)
        s = O.Sim(p, n)
        s.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"], V0=V0)
        t0 = time.perf_counter()
        s.substeps(sc["dt"], steps)
        secs = time.perf_counter() - t0
        return n * steps / secs, secs

    serial, serial_s = run(1, 2)
    parallel, parallel_s = run(threads, 4) if threads > 1 else (serial, serial_s)
    out = dict(value=serial, unit="particle-substeps/s", cores=1, kind="port",
               sample=f"2 substeps of the same {n}-particle slab on the 256^3 grid, oracle FP32, "
                      f"1 thread, {serial_s:.1f} s",
               all_cores=dict(value=parallel, cores=threads, nproc=os.cpu_count(),
                              sample=f"4 substeps, {threads} OpenMP threads, {parallel_s:.1f} s",
                              speedup_over_serial=parallel / serial))
    cal_path = os.path.join(ROOT, "oracle", "calibration.json")
    if os.path.exists(cal_path):
        with open(cal_path) as f:
            cal = json.load(f)
        r = cal["oracle_over_reference"]
        out["calibration"] = dict(oracle_over_reference_per_core=r, reference_per_core=cal["reference_per_core"],
                                  workload=cal["workload"], measured_on=cal["cpu"],
                                  reference_equivalent_serial=serial / r,
                                  reference_equivalent_all_cores=parallel / r)
    return out


def aligned_warmup(sim, dt, warmup):
    """Runs at least `warmup` substeps and stops on a re-bin boundary, so the
    timed window that follows starts with a re-bin and holds ceil(K / r)
    re-bins for K substeps at interval r (whole cycles are charged, never
    fewer); returns the substeps run."""
    r, w = sim.tuning()
    sim.set_tuning(r, w)  # the next substep re-bins: a cycle starts there
    w = max(r, (warmup + r - 1) // r * r)
    sim.substeps(dt, w)
    sim.sync()
    return w


def timed_substeps(sim, dt, steps):
    """K substeps as the library runs them (HIP-graph segments), wall clock
    between stream synchronisations, then K more launched directly with HIP
    events around every launch, from the start of a re-bin cycle; ->
    (wall_ms, stream_event_ms, direct_ms, per-kernel {name: (ms, launches)})."""
    sim.sync()
    t0 = time.perf_counter()
    sim.profile_begin(per_kernel=False)
    sim.substeps(dt, steps)
    gpu_ms, _ = sim.profile_end()
    sim.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3
    # the per-kernel pass starts a re-bin cycle too, so the binning's own
    # time is measured beside the other kernels
    r, w = sim.tuning()
    sim.set_tuning(r, w)
    sim.profile_begin(per_kernel=True)
    sim.substeps(dt, steps)
    direct_ms, kern = sim.profile_end()
    sim.sync()
    return wall_ms, gpu_ms, direct_ms, kern


def reproducible_timing(sim, dt, steps, n):
    """The same window in reproducible mode (mpm_set_reproducible: P2G
    contributions rounded to the scene's quanta, FP64 grid sums that repeat
    bit for bit), from the start of a re-bin cycle; the mode is switched
    off again afterwards."""
    sim.set_reproducible(True)
    aligned_warmup(sim, dt, 1)
    wall_ms, _, _, kern = timed_substeps(sim, dt, steps)
    sim.set_reproducible(False)
    return {"value": n * steps / (wall_ms / 1e3), "ms_per_step": wall_ms / steps,
            "kernels_ms": {k: (v[0] / max(1, v[1])) for k, v in kern.items() if v[1]}}


def roofline_of(kern, n, a_over_n, workload_key):
    dom = max(kern, key=lambda k: kern[k][0])
    dom_ms = kern[dom][0] / max(1, kern[dom][1])
    dom_bytes = alg_bytes_per_particle(dom, a_over_n) * n
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": _profiled_traffic(dom, workload_key),
            "alg_bytes_per_particle": alg_bytes_per_particle(dom, a_over_n), "mean_launch_ms": dom_ms}
    issue = None
    insts = _profiled_valu(dom, workload_key)
    if insts is not None:
        # vector instructions (64 lanes each) per launch, measured by
        # rocprofv3, over this run's mean launch time
        lane_ops = insts * 64 / (dom_ms / 1e3)
        issue = {"kernel": dom, "unit": "VALU lane-ops/s", "valu_insts_per_launch": insts, "achieved": lane_ops,
                 "peak": VALU_PEAK_LANE_OPS, "frac": lane_ops / VALU_PEAK_LANE_OPS}
        roof["measured_limiter"] = ("VALU/LDS instruction issue (issue_ceiling.frac %.2f of the VALU peak against "
                                    "frac %.2f of HBM)" % (issue["frac"], roof["frac"]))
    return dom, dom_ms, roof, issue


def deformation_summary(P, theta_c, theta_s, sample=200000, seed=0):
    """How far the particles' elastic state is from the identity, on a
    random sample: fraction with max |F_E - I| > 1e-4, fraction with a
    singular value on a clamp bound, mean J_P."""
    import numpy as np

    rng = np.random.default_rng(seed)
    idx = rng.choice(P["FE"].shape[0], size=min(sample, P["FE"].shape[0]), replace=False)
    FE = P["FE"][idx].reshape(-1, 3, 3).astype(np.float64)
    FP = P["FP"][idx].reshape(-1, 3, 3).astype(np.float64)
    dev = np.abs(FE - np.eye(3)).max(axis=(1, 2))
    sv = np.linalg.svd(FE, compute_uv=False)
    lo, hi = float(np.float32(1 - theta_c)), float(np.float32(1 + theta_s))
    at_clamp = ((np.abs(sv - lo) < 2e-6) | (np.abs(sv - hi) < 2e-6)).any(axis=1)
    return {"sample": int(idx.size), "frac_FE_off_identity_1e-4": float((dev > 1e-4).mean()),
            "frac_at_clamp": float(at_clamp.mean()), "max_FE_minus_I": float(dev.max()),
            "mean_JP": float(np.linalg.det(FP).mean())}


def bench_landing(M, args, steps, warmup):
    """The deforming counterpart of the headline: the same 4,194,304-particle
    slab landing on the sticky floor at 2 m/s (scenes.landing_slab), timed
    after at least 200 untimed substeps of impact (the autotuner runs after
    the first 100, as on the headline)."""
    from mpm_amd import scenes

    sc = scenes.landing_slab()
    n = sc["x"].shape[0]
    prm = sc["params"]
    p = M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_LITERAL, bc_mode=prm["bc_mode"], E=prm["E"],
                         nu=prm["nu"], xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"])
    sim = M.Simulation(p, n)
    sim.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    sim.init_volumes()
    dt = sc["dt"]
    sim.substeps(dt, 100)  # the first of the impact, untimed
    sim.sync()
    tuned = None
    if not args.no_autotune:
        # from the impact state; restored (unbinned) afterwards
        tuned = dict(ms_per_substep=sim.autotune(dt, 128))
        tuned["resort_every"], tuned["block_grid"] = sim.tuning()
    r, _ = sim.tuning()
    # more impact, untimed, ending on a re-bin boundary
    impact = 100 + aligned_warmup(sim, dt, max(100, warmup))
    wall_ms, gpu_ms, direct_ms, kern = timed_substeps(sim, dt, steps)
    a_over_n = sim.count_active() / n
    dom, dom_ms, roof, issue = roofline_of(kern, n, a_over_n, "slab4m_landing")
    status = sim.scene_status()
    P = sim.particles()
    sim.close()
    out = {"workload": "the same slab landing on the sticky floor at 2 m/s, timed after %d substeps of impact"
                       % impact,
           "value": n * steps / (wall_ms / 1e3), "ms_per_step": wall_ms / steps, "steps": steps,
           "resort_every": r, "rebins_timed": -(-steps // r), "roofline": roof,
           "kernels_ms": {k: (v[0] / max(1, v[1])) for k, v in kern.items() if v[1]},
           "deformation": deformation_summary(P, prm["theta_c"], prm["theta_s"]), "status": status}
    if issue:
        out["issue_ceiling"] = issue
    if tuned:
        out["autotune"] = tuned
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed substeps (default: 200 for slab4m and adjoint16, 500 for batch64)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-cores CPU baseline (0: host share)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device-to-device copy (calibrates rocprof byte counters)")
    ap.add_argument("--workload", default="slab4m", choices=["slab4m", "slab4m_landing", "batch64", "adjoint16"],
                    help="slab4m: the headline 4M-particle slab (with the landing slab beside it); "
                         "slab4m_landing: only the deforming landing slab; batch64: 64 independent scenes in one "
                         "context against the same scenes run one after another; adjoint16: forward and "
                         "reverse mode over 16 scenes of 32,768 particles")
    ap.add_argument("--autotune", action="store_true",
                    help="batch64: run the library's autotuner (re-bin interval, block-pass width) before timing")
    ap.add_argument("--no-autotune", action="store_true",
                    help="slab4m: time the library's default tuning instead of the autotuner's choice")
    ap.add_argument("--no-landing", action="store_true", help="slab4m: skip the deforming landing slab")
    ap.add_argument("--no-secondary", action="store_true",
                    help="slab4m: skip the batch64 and adjoint16 records (each is also its own --workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); "
              "multi-GPU sharding of the substep is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 500 if args.workload == "batch64" else 200
    if args.workload == "batch64":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "adjoint16":
        print(json.dumps(bench_adjoint(args)), flush=True)
        return

    import mpm_amd as M
    from mpm_amd import scenes

    if args.workload == "slab4m_landing":
        print(json.dumps(bench_landing(M, args, args.steps, args.warmup)), flush=True)
        return

    t0 = time.perf_counter()
    sc = scenes.slab()
    host_gen_ms = (time.perf_counter() - t0) * 1e3
    n = sc["x"].shape[0]
    prm = sc["params"]
    p = M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_LITERAL, bc_mode=prm["bc_mode"], E=prm["E"],
                         nu=prm["nu"], xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"])
    sim = M.Simulation(p, n)
    t0 = time.perf_counter()
    sim.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    upload_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    sim.init_volumes()
    sim.sync()
    volumes_ms = (time.perf_counter() - t0) * 1e3
    V0 = sim.particles()["V0"] if not args.no_cpu_baseline else None
    dt = sc["dt"]
    tuned = None
    if not args.no_autotune:
        # mpm_autotune times each candidate from the current state and
        # restores it (unbinned); the timed substeps below start from it
        tuned = dict(ms_per_substep=sim.autotune(dt, 128))
        tuned["resort_every"], tuned["block_grid"] = sim.tuning()
    resort, _ = sim.tuning()
    warm_run = aligned_warmup(sim, dt, args.warmup)
    wall_ms, gpu_ms, direct_ms, kern = timed_substeps(sim, dt, args.steps)
    _note(f"slab4m: {args.steps} substeps timed")
    active = sim.count_active()
    a_over_n = active / n
    ms_per_step = wall_ms / args.steps
    value = n * args.steps / (wall_ms / 1e3)
    dom, dom_ms, roof, issue = roofline_of(kern, n, a_over_n, "slab4m")
    sub_bytes = substep_alg_bytes(a_over_n) * n
    rebins = -(-args.steps // resort)
    bin_ms = kern.get("bin", (0.0, 0))
    bin_mean = bin_ms[0] / max(1, bin_ms[1])
    out = {
        "metric": "particle-substeps/s at 4M particles (% HBM roofline); fwd+bwd substeps/s",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "snow slab 4,194,304 particles, 256^3 grid, 4^3 blocks, dt 2.5e-5, E0 U[1e5,2e5]",
                   "particles": n, "grid": list(sc["dims"]), "active_nodes": active, "compose": "literal"},
        "roofline": roof,
        "substep_roofline": {"alg_bytes_per_particle": substep_alg_bytes(a_over_n),
                             "achieved": sub_bytes / (ms_per_step / 1e3) / 1e9,
                             "frac": sub_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernels_ms": {k: (v[0] / max(1, v[1])) for k, v in kern.items() if v[1]},
        "timing": {"wall_ms": wall_ms, "stream_event_ms": gpu_ms, "direct_launch_ms": direct_ms,
                   "warmup_substeps_run": warm_run, "resort_every": resort, "rebins_timed": rebins,
                   # the same window with the re-bin cost spread at 1 / resort_every per substep
                   "ms_per_step_rebin_amortised": (wall_ms - (rebins - args.steps / resort) * bin_mean) / args.steps},
    }
    if issue:
        out["issue_ceiling"] = issue
    if tuned:
        out["autotune"] = tuned
    out["reproducible"] = reproducible_timing(sim, dt, args.steps, n)
    out["setup_io"] = setup_and_io(M, sim, sc, dt, host_gen_ms, upload_ms, volumes_ms)
    if args.copy_probe:
        sim.copy_probe(1 << 30)
    sim.close()
    _note("slab4m: reproducible mode and setup/readback timed")
    if not args.no_landing:
        out["deforming"] = bench_landing(M, args, args.steps, args.warmup)
        _note("slab4m_landing: timed")
    if not args.no_secondary:
        # the two secondary workloads at their own sizes and substep counts
        out["batch64"] = secondary(bench_batch, 500, args.warmup)
        _note("batch64: timed")
        out["adjoint16"] = secondary(bench_adjoint, 200, args.warmup)
        _note("adjoint16: timed")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, V0, args.cpu_threads or _host_threads())
        _note("cpu_baseline: timed")
    print(json.dumps(out), flush=True)


def setup_and_io(M, sim, sc, dt, host_gen_ms, upload_ms, volumes_ms):
    """Setup and readback around the timed region (not part of `value`):
    host generation + upload against the device generator on the same-size
    slab, and the viewer's positions readback (synchronous, and begun before
    10 substeps so the PCIe copy overlaps them).  Leaves the context holding
    the device-generated slab."""
    n = sim.n
    io = {"host_generate_ms": host_gen_ms, "host_upload_ms": upload_ms, "init_volumes_ms": volumes_ms}
    sim.positions()  # allocates the staging buffer and pinned mirror
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.positions()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    io["positions_readback_ms"] = ms
    io["positions_readback_GBps"] = 12.0 * n / (ms * 1e6)
    sim.sync()
    t0 = time.perf_counter()
    sim.substeps(dt, 10)
    sim.sync()
    io["substeps10_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    sim.positions_begin()
    sim.substeps(dt, 10)
    sim.positions_wait()
    sim.sync()
    io["substeps10_with_async_readback_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    sim.generate(M.GEN_CELLS, seed=2, cell_lo=(64, 128, 64), cells=(128, 32, 128), ppc=(4, 12),
                 mass=float(sc["m"][0]), E=(1e5, 2e5))
    sim.sync()
    io["device_generate_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    sim.init_volumes()
    sim.sync()
    io["device_init_volumes_ms"] = (time.perf_counter() - t0) * 1e3
    return io


def _scene_contact(P, dims, h):
    """True when some particle's stencil reaches the domain's wall bands
    (node rows < 3 or >= dims - 3): x / h < 4 or > dims - 4 on an axis."""
    import numpy as np

    q = P["x"] / h
    return bool(((q < 4.0) | (q > np.asarray(dims) - 4.0)).any())


def bench_batch(args):
    """64 scenes (N_s log-uniform in [4096, 262144], 64^3 or 128^3 grids,
    E0 log-uniform in [5e4, 5e5], v0 uniform in [-3, 3]^3, dt 1e-4) for the
    configuration's 500 substeps: one batched context against one context
    per scene run one after another, each timed over the same substeps from
    the same state, re-bins included.  Reports how many scenes reached a wall
    band and how far F_E left the identity by the end."""
    import numpy as np
    import mpm_amd as M
    from mpm_amd import scenes

    scs = scenes.scene_batch()
    ns = [s["x"].shape[0] for s in scs]
    dt = scs[0]["dt"]
    k = args.steps

    def params(sc):
        prm = sc["params"]
        return M.default_params(sc["dims"], sc["h"], bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"], xi=prm["xi"],
                                theta_c=prm["theta_c"], theta_s=prm["theta_s"])

    b = M.Simulation([params(s) for s in scs], ns)
    for i, s in enumerate(scs):
        b.scene_set_particles(i, s["x"], s["v"], s["m"], s["E0"])
    b.init_volumes()
    tuned = None
    if args.autotune:
        tuned = dict(ms_per_substep=b.autotune(dt, 16))
        tuned["resort_every"], tuned["block_grid"] = b.tuning()
    warm = aligned_warmup(b, dt, args.warmup)
    t0 = time.perf_counter()
    b.substeps(dt, k)
    b.sync()
    batch_s = time.perf_counter() - t0
    status = b.scene_status()
    parts = [b.scene_particles(i) for i in range(len(scs))]
    b.close()
    contact = [_scene_contact(P, s["dims"], s["h"]) for P, s in zip(parts, scs)]
    allP = {"FE": np.concatenate([P["FE"] for P in parts]), "FP": np.concatenate([P["FP"] for P in parts])}
    prm0 = scs[0]["params"]

    singles = []
    for s in scs:
        g = M.Simulation(params(s), s["x"].shape[0])
        g.set_particles(s["x"], s["v"], s["m"], s["E0"])
        g.init_volumes()
        aligned_warmup(g, dt, args.warmup)
        singles.append(g)
    t0 = time.perf_counter()
    for g in singles:
        g.substeps(dt, k)
        g.sync()
    seq_s = time.perf_counter() - t0
    for g in singles:
        g.close()

    total = int(sum(ns))
    value = total * k / batch_s
    out = {
        "metric": "batched particle-substeps/s, 64 scenes (speed-up vs one scene at a time)",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": batch_s / k * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 snowball scenes, N_s log-U[4096, 262144], 64^3/128^3, dt 1e-4",
                   "particles": total, "scenes": len(scs), "grids_128": int(sum(1 for s in scs if s["dims"][0] == 128)),
                   "warmup_substeps_run": warm},
        "sequential": {"value": total * k / seq_s, "ms_per_step": seq_s / k * 1e3},
        "speedup_vs_sequential": seq_s / batch_s,
        "nonfinite_scenes": [i for i, st in enumerate(status) if st == M.MPM_ERR_NONFINITE],
        "cfl_warning_scenes": [i for i, st in enumerate(status) if st == M.MPM_ERR_CFL],
        "scenes_at_a_wall_at_end": int(sum(contact)),
        "deformation_at_end": deformation_summary(allP, prm0["theta_c"], prm0["theta_s"]),
    }
    if tuned:
        out["autotune"] = tuned
    return out


def bench_adjoint(args):
    """The differentiable workload: 16 compressing snowballs
    (scenes.adjoint_batch) of 32,768 particles on 64^3 (TEXTBOOK composition,
    domain walls, dt 1e-4): `steps` substeps forward, keeping checkpoints at
    the interval the library chooses (every substep when the records fit in
    device memory), then the reverse sweep of the loss |centre of mass -
    target|^2 + 1e-3 mean |v|^2; the same with checkpoints every 20 (the
    workload's stated interval) beside it.  A step is one forward plus one backward
    substep."""
    import numpy as np
    import mpm_amd as M

    from mpm_amd import scenes

    steps = args.steps
    scs, targets = scenes.adjoint_batch()
    ns = [s["x"].shape[0] for s in scs]
    dt = scs[0]["dt"]

    def params(sc):
        prm = sc["params"]
        return M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_TEXTBOOK, bc_mode=prm["bc_mode"], E=prm["E"],
                                nu=prm["nu"], xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"])

    b = M.Simulation([params(s) for s in scs], ns)

    def reset():
        for i, s in enumerate(scs):
            b.scene_set_particles(i, s["x"], s["v"], s["m"], s["E0"])
        b.init_volumes()
        b.sync()

    reset()
    n = b.n
    # the loss gradient at the final state (the same for every interval up
    # to rounding), then one warm-up of each timed configuration (the
    # library keeps its checkpoint, record and workspace buffers)
    b.adjoint_forward(dt, steps, 0)
    fin = [b.scene_particles(i) for i in range(len(scs))]
    loss, gxs, gvs = scenes.com_loss([P["x"] for P in fin], [P["v"] for P in fin], targets)
    gx, gv = np.concatenate(gxs), np.concatenate(gvs)
    b.adjoint_backward(gx, gv)
    b.adjoint_forward(dt, steps, 20)
    b.adjoint_backward(gx, gv)
    # the plain forward warmed up too: its segment graphs are captured on
    # first use
    reset()
    b.substeps(dt, steps)
    b.sync()
    reset()
    t0 = time.perf_counter()
    b.substeps(dt, steps)
    b.sync()
    plain_s = time.perf_counter() - t0
    status = b.scene_status()
    reset()

    def run(every):
        t0 = time.perf_counter()
        b.adjoint_forward(dt, steps, every)
        b.sync()
        f = time.perf_counter() - t0
        chosen = b.adjoint_interval()
        t0 = time.perf_counter()
        g = b.adjoint_backward(gx, gv)
        b.sync()
        return f, time.perf_counter() - t0, chosen, g

    fwd_s, bwd_s, chosen, (gx0, gv0, gE0) = run(0)
    fwd20_s, bwd20_s, _, (gx1, gv1, gE1) = run(20)
    b.close()

    def rel(a, r):
        return float(np.linalg.norm(a.astype(np.float64) - r) / max(np.linalg.norm(r.astype(np.float64)), 1e-30))
    off = np.r_[0, np.cumsum(ns)]
    dE = [float(gE0[off[i]:off[i + 1]].astype(np.float64).sum()) for i in range(len(scs))]
    dv = [gv0[off[i]:off[i + 1]].astype(np.float64).sum(0).tolist() for i in range(len(scs))]
    cells = sum(int(np.prod(sc["dims"])) for sc in scs)
    out = {
        "metric": "fwd+bwd substeps/s, 16 x 32,768 particles on 64^3 (checkpoint interval chosen by the library)",
        "value": steps / (fwd_s + bwd_s),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": (fwd_s + bwd_s) / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "16 compressing snowballs (1.2/s) x 32,768 particles, 64^3, dt 1e-4, TEXTBOOK, "
                               "loss |com - target|^2 + 1e-3 mean |v|^2",
                   "particles": n, "scenes": len(scs)},
        "loss": loss,
        "checkpoint_every": chosen,
        "forward_ms_per_substep": fwd_s / steps * 1e3,
        "backward_ms_per_substep": bwd_s / steps * 1e3,
        "plain_forward_ms_per_substep": plain_s / steps * 1e3,
        "fwd_plus_bwd_over_plain_fwd": (fwd_s + bwd_s) / plain_s,
        "particle_substeps_per_s_fwd_bwd": n * steps / (fwd_s + bwd_s),
        # state rows, room for every grid node's P2G sums, the block list
        "record_bytes_per_substep": 132 * n + 16 * cells + 4 * (cells // 64 + 1),
        "checkpoints_every_20": {
            "forward_ms_per_substep": fwd20_s / steps * 1e3,
            "backward_ms_per_substep": bwd20_s / steps * 1e3,
            "fwd_plus_bwd_over_plain_fwd": (fwd20_s + bwd20_s) / plain_s,
            "fwd_plus_bwd_over_checkpointed_fwd": (fwd20_s + bwd20_s) / fwd20_s,
            "grad_rel_diff": {"x0": rel(gx1, gx0), "v0": rel(gv1, gv0), "E0": rel(gE1, gE0)},
        },
        "dL_dE0_per_scene": dE,
        "dL_dv0_per_scene": dv,
        "status": status,
        "grad_finite": bool(np.isfinite(gx0).all() and np.isfinite(gv0).all() and np.isfinite(gE0).all()),
    }
    return out


_CONTRACT_ONLY = ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data")


def secondary(fn, steps, warmup):
    """A secondary workload's record inside the headline line: the same
    measurement as its own --workload run, without the keys the line
    already states."""
    rec = fn(argparse.Namespace(steps=steps, warmup=warmup, autotune=False))
    return {k: v for k, v in rec.items() if k not in _CONTRACT_ONLY}


def _profiled(name, workload, kernel, key):
    """One per-launch counter figure of `kernel` on `workload` from a
    committed rocprofv3 --pmc summary under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    entry = data.get("workloads", {}).get(workload, {}).get("kernels", {}).get(kernel)
    return None if entry is None else entry.get(key)


def _profiled_traffic(kernel, workload="slab4m"):
    """HBM bytes per launch (FETCH_SIZE and WRITE_SIZE passes scaled by a copy
    kernel of known size), profiles/traffic.json."""
    return _profiled("traffic.json", workload, kernel, "bytes_per_launch")


def _profiled_valu(kernel, workload="slab4m"):
    """Vector instructions per launch (SQ_INSTS_VALU), profiles/issue.json."""
    return _profiled("issue.json", workload, kernel, "SQ_INSTS_VALU")


if __name__ == "__main__":
    main()
