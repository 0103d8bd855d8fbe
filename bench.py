"""Benchmark: D3Q19 MRT Taylor-Green vortex, 256^3 cells, FP32 populations, one MI355X.

A step is one ``Simulation.one_time_step`` pass of the hot path over the
256^3 periodic box (for this fully periodic box the fused pull/collide kernel
reads through the periodic wrap, so a step is exactly one kernel launch).
The timed region starts with the populations resident in HBM: W warm-up steps,
device sync, K timed steps between HIP events on the simulation's stream,
device sync.  Printed as one JSON line:

  value     MLUPS = cells * K / seconds / 1e6 (pylbm/monitoring.py:111)
  roofline  algorithmic bytes of the fused kernel (19 loads + 19 stores of
            4 B per cell = 152 B/cell) per launch / average launch time,
            against the 8 TB/s HBM peak; traffic = HBM bytes per launch from
            rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/), scaled by the same
            counters on a plain copy kernel of known size, or null
  cpu_baseline  the oracle's C restatement of pylbm's generated Cython loop
            (FP64, AoS, -O3) on a bounded 128^3 sample, all host threads
  configs   the C1 cavity (FP64, 2000 steps), C2 channel (20000 steps), C4 sphere
            packing (5000 steps) and the C5 plain lattice run (2000 steps): MLUPS,
            the FP32 runs' moment errors against the FP64 run on the GPU, each link
            list timed alone, mass drifts
  surrogate the learned surrogate on the 512^2 periodic D2Q9 decaying flow:
            trained on HIP lattice trajectories, then its simulated steps per
            second against the plain FP32 lattice run (the temporally blocked run,
            and one fused launch per step) and its velocity rel-L2

``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + % HBM roofline (D3Q19 256³ FP32); surrogate speedup at ≤1e-2 rel-L2"
HBM_PEAK_GBPS = 8000.0
BYTES = {"float32": 4, "float64": 8, "float16": 2, "bfloat16": 2}
ERR_STEPS = 1000   # steps of the accuracy leg (the D3Q19 256^3 config states its error after 1000)
VARIANT_WARMUP = 100   # untimed steps before each variant's 200 timed steps
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fused_step_d3q19_256_fp32.txt")


def _time_run(sim, steps, fold=None):
    rt = sim.rt
    e0, e1 = rt.event(), rt.event()
    rt.sync()
    t0 = time.perf_counter()
    rt.record(e0)
    sim.run(steps, fold=fold)
    rt.record(e1)
    ms = rt.elapsed_ms(e0, e1)
    rt.sync()
    wall = time.perf_counter() - t0
    return ms, wall


def _copy_bandwidth(rt, nbytes=2 << 30, reps=20):
    """GB/s of the plain copy kernel (read + write) over nbytes buffers."""
    src, dst = rt.malloc(nbytes), rt.malloc(nbytes)
    try:
        rt.check(rt.lib.lbm_copy(dst, src, nbytes, rt.stream), "lbm_copy")
        e0, e1 = rt.event(), rt.event()
        rt.record(e0)
        for _ in range(reps):
            rt.check(rt.lib.lbm_copy(dst, src, nbytes, rt.stream), "lbm_copy")
        rt.record(e1)
        ms = rt.elapsed_ms(e0, e1)
    finally:
        rt.free(src)
        rt.free(dst)
    return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9


def _traffic():
    """HBM bytes per fused-kernel launch from the committed rocprofv3 PMC summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    txt = open(TRAFFIC_FILE).read()
    m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def _cpu_baseline(target_s=8.0):
    """MLUPS of the C restatement of the generated Cython loop on a 128^3 sample."""
    from oracle.cgen import CStepper, set_threads
    from oracle.scheme import Scheme as OScheme
    from pylbm_amd import cases

    n = 128
    d = cases.taylor_green_d3q19(n)
    sch = OScheme(d)
    # initial state: equilibrium of the Taylor-Green fields (same init as the GPU run)
    xs = np.arange(n + 2) - 0.5
    X, Y, Z = np.meshgrid(xs, xs, xs, sparse=True, indexing="ij")
    init = d["init"]
    from pylbm_amd.cases import QX, QY, RHO
    mc = [np.broadcast_to(init[RHO](X, Y, Z), (n + 2,) * 3), np.broadcast_to(init[QX](X, Y, Z), (n + 2,) * 3),
          np.broadcast_to(init[QY](X, Y, Z), (n + 2,) * 3), np.zeros((n + 2,) * 3)]
    eq = np.stack(sch.equilibrium(mc))
    f = np.ascontiguousarray(np.moveaxis(np.tensordot(sch.invMnum, eq, axes=(1, 0)), 0, -1))
    g = np.empty_like(f)
    cs = CStepper(sch)
    halo = (1, 1, 1)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    # every core of the process's CPU share; a thread limit set for the process
    # (OMP_NUM_THREADS) is that share where the machine is shared between jobs
    limit = os.environ.get("OMP_NUM_THREADS")
    if limit and limit.isdigit() and int(limit) > 0:
        cores = min(cores, int(limit))
    out = {}
    for threads in (1, cores):
        set_threads(threads)
        t0 = time.perf_counter()
        cs.ghost(f, halo)
        cs.step(f, g, halo, 0.0, 1.0)
        one = time.perf_counter() - t0
        steps = max(1, int(target_s / max(one, 1e-3)))
        t0 = time.perf_counter()
        for _ in range(steps):
            cs.ghost(f, halo)
            cs.step(f, g, halo, 0.0, 1.0)
            f, g = g, f
        sec = time.perf_counter() - t0
        out[threads] = (n ** 3 * steps / sec / 1e6, steps)
    return out, cores, n


def _host_info():
    """CPU model, logical CPUs of the machine and of this process."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _link_timing(sim, reps=20):
    """Per-launch time and achieved bytes of each link list of sim, applied alone (a list
    fused into the step, fused_into_step, is timed as the launch it replaces)."""
    rt = sim.rt
    out = []
    F = sim.container.F
    es = F.layout.elem_size
    for dl in sim._dev_links:
        if dl.n == 0:
            continue
        dl.apply(F.ptr)
        e0, e1 = rt.event(), rt.event()
        rt.record(e0)
        for _ in range(reps):
            dl.apply(F.ptr)
        rt.record(e1)
        ms = rt.elapsed_ms(e0, e1) / reps
        cs = 8 if es == 8 else 4   # compute type of rhs and coefficients
        rhs_bytes = cs if dl.d_rhs is not None else 0
        if getattr(dl, "bouzidi", False):
            # three indices, two loads and a store, two coefficients
            per_link = 12 + 3 * es + 2 * cs + rhs_bytes
        else:
            per_link = 8 + 2 * es + rhs_bytes + (1 if dl.d_sign is not None else 0)
        name = getattr(dl.method, "name", None) or str(getattr(dl.method, "kind", dl.method))
        out.append({"method": str(name), "links": int(dl.n), "fused_into_step": bool(getattr(dl, "fused", False)),
                    "us_per_launch": round(1e3 * ms, 2),
                    "Glinks_per_s": round(dl.n / (ms * 1e-3) / 1e9, 3),
                    "GBps_alg": round(dl.n * per_link / (ms * 1e-3) / 1e9, 1), "bytes_per_link": per_link})
    return out


def _moment_err(sim, ref, syms):
    """Normwise max relative error per conserved moment on the scale max(||m_ref||, 1e-3 ||rho_ref||)."""
    mref = [ref.m[x] for x in syms]
    floor = 1e-3 * float(np.max(np.abs(mref[0])))
    return max(float(np.max(np.abs(sim.m[x] - b))) / max(float(np.max(np.abs(b))), floor)
               for x, b in zip(syms, mref))


def _configs_leg(only=None, steps_override=None):
    """The other workloads at their stated run lengths: MLUPS through Simulation.run, each
    FP32 run's moment error against the same run in FP64 on the GPU, each link list
    timed alone, and mass drifts where the boundaries conserve mass.  only: run just the
    configurations whose name contains it; steps_override: that many steps instead of
    the stated run lengths (both for a profiler)."""
    from pylbm_amd import Simulation, cases

    res = {}
    r2, r3 = (cases.RHO, cases.QX, cases.QY), (cases.RHO, cases.QX, cases.QY, cases.QZ)
    plan = [("C1 lid-driven cavity 128^2 D2Q9 MRT", lambda: cases.lid_cavity_d2q9(128), ("float64",), 2000, r2, False),
            ("C2 Karman channel 1024x256 D2Q9 MRT", lambda: cases.karman_d2q9(), ("float32", "float32-shifted", "float64"),
             20000, r2, False),
            ("C2 Karman channel 1024x256 D2Q9 MRT, Bouzidi BB/ABB", lambda: cases.with_bouzidi(cases.karman_d2q9()),
             ("float32-shifted", "float64"), 20000, r2, False),
            ("C4 sphere packing 512x256x256 D3Q19 MRT", lambda: cases.sphere_packing_d3q19(),
             ("float32", "float32-shifted", "float64"), 5000, r3, True),
            ("C5 decaying flow 512^2 D2Q9 MRT (plain lattice run)", lambda: cases.decaying_d2q9(512),
             ("float32", "float64"), 2000, r2, True)]
    for name, make, dtypes, steps, syms, mass in plan:
        if only and only not in name:
            continue
        steps = steps_override or steps
        entry = {"steps": steps}
        sims = {}
        for dt in dtypes:
            t0 = time.perf_counter()
            sim = Simulation(make(), dtype=dt)
            setup = time.perf_counter() - t0
            cells = int(np.prod(sim.domain.shape_in))
            m0 = sim.conserved_sums()[0] if mass else None
            ms, _ = _time_run(sim, steps)
            e = {"MLUPS": round(cells * steps / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms / steps, 4),
                 "setup_s": round(setup, 1)}
            if mass:
                e["mass_rel_drift"] = float("%.3e" % ((sim.conserved_sums()[0] - m0) / m0))
            entry[dt] = e
            entry["cells"] = cells
            if sim.domain.elements and "solid_fraction" not in entry:
                inner = tuple(slice(h, -h) if h else slice(None) for h in sim.domain.stencil.vmax)
                entry["solid_fraction"] = round(1.0 - float(np.mean(sim.domain.fluid[inner])), 5)
            sims[dt] = sim
            print("bench.py: %s %s: %.4f ms per step" % (name, dt, ms / steps), file=sys.stderr, flush=True)
        for dt in dtypes:
            if dt != "float64" and "float64" in sims:
                entry[dt]["moment_err_vs_float64"] = float("%.3e" % _moment_err(sims[dt], sims["float64"], syms))
        # after the errors: timing the lists alone rewrites their store slots (solid cells, halo)
        first = sims[dtypes[0]]
        entry["links"] = _link_timing(first)
        res[name] = entry
        del sims, first
    return res


def _surrogate_leg(ntraj, steps, steps_per_traj, eval_seeds=(1000001, 1000002), K=16, judge_every=8,
                   tol=1e-2, horizons=(256, 2000), elem="float16"):
    """Train the surrogate on HIP lattice trajectories (512^2 periodic D2Q9, FP32, nu = 1e-3),
    then advance held-out initial states by each horizon with the plain run and with the
    judged surrogate: simulated steps per second of both (wall clock, device-resident
    state, one state read-back in both), the surrogate's velocity rel-L2 against the
    plain run, and the same error for the plain run restarted from its own moments every
    K steps (the floor of any method that carries only the moments)."""
    from pylbm_amd.surrogate import LatticeState, SurrogateRun, train_surrogate

    lat = LatticeState(n=512, dtype="float32")
    rt = lat.rt
    net, norm, info = train_surrogate(lat, ntraj=ntraj, steps=steps, K=K, steps_per_traj=steps_per_traj, elem=elem)
    run = SurrogateRun(net, lat, norm, K=K, judge_every=judge_every, tol=tol)
    ref = rt.malloc(lat.npix * 16)
    work = rt.malloc(lat.npix * 16)
    # one network step (3 hidden convolutions + the fused output/apply convolution):
    # HIP events around 200 steps of the C loop on the lattice stream
    ch = (64,) + (info["hidden"],) * 3 + (64,)
    flops = 2 * 9 * net.H * net.W * sum(ch[i] * ch[i + 1] for i in range(4))
    rt.h2d(work, lat.initial_state(eval_seeds[0]))
    run._make_input(work)
    net.run(work, (run._dscale, run._shift, run._scale), 10)
    e0, e1 = rt.event(), rt.event()
    rt.record(e0)
    net.run(work, (run._dscale, run._shift, run._scale), 200)
    rt.record(e1)
    step_us = 1e3 * rt.elapsed_ms(e0, e1) / 200
    rt.record(e0)
    lat.run(16 * 25)
    rt.record(e1)
    lbm_step_us = 1e3 * rt.elapsed_ms(e0, e1) / 400
    lat.blocked = False
    lat.run(16)
    rt.record(e0)
    lat.run(16 * 25)
    rt.record(e1)
    lbm_step_us_one_launch = 1e3 * rt.elapsed_ms(e0, e1) / 400
    lat.blocked = True
    by_h = {}

    def judged(hz, je):
        """Per held-out state, over hz steps with a judge every je intervals: plain and
        judged-surrogate wall times, speedup, velocity rel-L2, fall-back, drifts."""
        nint = hz // K
        out = []
        run.judge_every = je
        for seed in eval_seeds:
            s0 = lat.initial_state(seed)
            rt.h2d(run.state, s0)
            lat.restart(run.state)
            lat.run(K)                       # warm-up of the plain path
            lat.restart(run.state)
            rt.sync()
            t0 = time.perf_counter()
            lat.run(K * nint)
            lat.read_state(ref)
            rt.sync()
            t_lbm = time.perf_counter() - t0
            # the same lattice run with one fused launch per step (the comparison
            # before the blocked run existed): same bits, more time
            lat.blocked = False
            lat.restart(run.state)
            rt.sync()
            t0 = time.perf_counter()
            lat.run(K * nint)
            lat.read_state(work)
            rt.sync()
            t_lbm1 = time.perf_counter() - t0
            lat.blocked = True
            rt.h2d(run.state, s0)
            lat.restart(run.state)
            rt.sync()
            t0 = time.perf_counter()
            run.start_from_lattice()
            run.advance(nint)
            rt.sync()
            t_sur = time.perf_counter() - t0
            err = run.rel_l2(run.state, ref)
            out.append({"seed": seed, "lbm_s": round(t_lbm, 5), "lbm_one_launch_per_step_s": round(t_lbm1, 5),
                        "surrogate_s": round(t_sur, 5), "speedup": round(t_lbm / t_sur, 3),
                        "speedup_vs_one_launch_per_step": round(t_lbm1 / t_sur, 3),
                        "rel_l2_velocity": float("%.3e" % err),
                        "fallback_at_interval": run.fallback_at,
                        "judge_drifts": [float("%.2e" % j) for j in run.judges]})
        run.judge_every = judge_every
        return out

    for hz in horizons:
        nint = hz // K
        evals = judged(hz, judge_every)
        # moments-only floor: the plain run restarted from its moments every K steps
        rt.h2d(work, lat.initial_state(eval_seeds[0]))
        lat.restart(work)
        for _ in range(nint):
            lat.run(K)
            lat.read_state(work)
            lat.restart(work)
        lat.read_state(work)
        rt.h2d(run.state, lat.initial_state(eval_seeds[0]))
        lat.restart(run.state)
        lat.run(K * nint)
        lat.read_state(ref)
        floor = run.rel_l2(work, ref)
        by_h[str(hz)] = {"speedup": round(min(e["speedup"] for e in evals), 3),
                         "speedup_vs_one_launch_per_step": round(min(e["speedup_vs_one_launch_per_step"]
                                                                     for e in evals), 3),
                         "rel_l2_velocity": max(e["rel_l2_velocity"] for e in evals),
                         "moments_only_floor_rel_l2": float("%.3e" % floor), "eval": evals}
    # sensitivity (not the configured schedule): the first horizon judged every 16 intervals
    ev16 = judged(horizons[0], 16)
    sens = {"horizon": horizons[0], "judge_every": 16, "speedup": round(min(e["speedup"] for e in ev16), 3),
            "rel_l2_velocity": max(e["rel_l2_velocity"] for e in ev16), "eval": ev16}
    rt.free(ref)
    rt.free(work)
    run.free()
    net.free()
    info.pop("losses", None)
    return {
        "workload": "512^2 periodic D2Q9 MRT decaying flow, FP32, nu=1e-3, random Fourier velocity "
                    "E(k)~k^-5/3 on |k| in [4,16], rms 0.05",
        "network": "4x4-patch periodic 3x3 conv net 64-%d-%d-%d-64, SiLU, %s MFMA, FP32 accumulate"
                   % ((info["hidden"],) * 3 + ({"float16": "FP16", "bfloat16": "BF16"}[elem],)),
        "K": K, "judge_every": judge_every, "tol_rel_l2": tol,
        "network_step_us": round(step_us, 2), "lattice_step_us": round(lbm_step_us, 3),
        "lattice_step_us_one_launch_per_step": round(lbm_step_us_one_launch, 3),
        "network_roofline": {"bound": "mfma", "achieved": round(flops / step_us * 1e-6, 1),
                             "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": round(flops / step_us * 1e-6 / 2500.0, 4),
                             "flops_per_step": flops,
                             "dtype": "%s x %s -> fp32" % (({"float16": "fp16", "bfloat16": "bf16"}[elem],) * 2)},
        "horizons": by_h,
        "judge_every_16_sensitivity": sens,
        "train": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in info.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the C1/C2/C4 timings")
    ap.add_argument("--config", default=None, help="time only the configurations whose name contains this")
    ap.add_argument("--config-steps", type=int, default=None,
                    help="steps per configuration run instead of the stated lengths (profiling)")
    ap.add_argument("--sur-traj", type=int, default=200, help="surrogate training trajectories")
    ap.add_argument("--sur-steps", type=int, default=2000, help="lattice steps per trajectory")
    ap.add_argument("--sur-train", type=int, default=100, help="optimizer steps per trajectory")
    ap.add_argument("--sur-elem", default="float16", choices=("float16", "bfloat16"),
                    help="16-bit MFMA operand type of the surrogate network")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="launch the plain copy kernel on a buffer of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from pylbm_amd import Simulation, cases

    n = args.n
    cells = n ** 3
    sim = Simulation(cases.taylor_green_d3q19(n), dtype="float32")
    rt = sim.rt
    if args.copy_probe:
        nb = args.copy_probe << 20
        src, dst = rt.malloc(nb), rt.malloc(nb)
        rt.check(rt.lib.lbm_copy(dst, src, nb, rt.stream), "lbm_copy")
        rt.sync()
        rt.free(src)
        rt.free(dst)
    sim.run(args.warmup)
    ms, wall = _time_run(sim, args.steps)
    sec = ms * 1e-3
    mlups = cells * args.steps / sec / 1e6
    alg_bytes = 2 * 19 * BYTES["float32"] * cells          # per launch
    achieved = alg_bytes * args.steps / sec / 1e9           # GB/s
    traffic = _traffic()
    print("bench.py: %d^3 FP32: %.4f ms per step" % (n, ms / args.steps), file=sys.stderr, flush=True)

    extra = {}
    if not args.no_variants:
        extra["copy_kernel_GBps"] = round(_copy_bandwidth(rt), 1)
        # moment error of every format against the FP64 run on the GPU after ERR_STEPS
        # steps (fixed, whatever --steps is), normwise per conserved moment on the scale
        # max(||m64||, 1e-3 ||rho64||); then each format's rate over 200 steps
        syms = (cases.RHO, cases.QX, cases.QY, cases.QZ)
        ref64 = Simulation(cases.taylor_green_d3q19(n), dtype="float64")
        ref64.run(ERR_STEPS)
        m64 = [ref64.m[s] for s in syms]
        floor = 1e-3 * float(np.max(np.abs(m64[0])))

        def moment_err(v):
            return max(float(np.max(np.abs(v.m[s] - b))) / max(float(np.max(np.abs(b))), floor)
                       for s, b in zip(syms, m64))

        errs = {}
        vmlups = {}
        for dt in ("float32", "float64", "float32-shifted", "float16", "bfloat16"):
            v = ref64 if dt == "float64" else Simulation(cases.taylor_green_d3q19(n), dtype=dt)
            if dt != "float64":
                v.run(ERR_STEPS)
                errs[dt] = moment_err(v)
            # untimed warm-up after the host-side error computation: the GPU ramps
            # back up over the first steps after an idle host pause
            v.run(VARIANT_WARMUP)
            vms, _ = _time_run(v, 200)
            r = cells * 200 / (vms * 1e-3) / 1e6
            vmlups[dt] = {"MLUPS": round(r, 1),
                          "frac_of_hbm_peak": round(r * 1e6 * 2 * 19 * BYTES[dt.split("-")[0]] / 1e9
                                                    / HBM_PEAK_GBPS, 4)}
            del v
        sim.run(VARIANT_WARMUP, fold=False)
        hms, _ = _time_run(sim, 200, fold=False)
        r = cells * 200 / (hms * 1e-3) / 1e6
        vmlups["float32_halo_ghost_fill"] = {"MLUPS": round(r, 1),
                                             "frac_of_hbm_peak": round(r * 1e6 * 152 / 1e9 / HBM_PEAK_GBPS, 4)}
        extra["variant_MLUPS"] = vmlups
        extra["moment_err_vs_fp64_after_%d_steps" % ERR_STEPS] = {k: float("%.3e" % v) for k, v in errs.items()}
        extra["moment_err_bar"] = {"float32": 1e-5, "met": errs["float32"] <= 1e-5}
        del ref64

    cpu = None
    if not args.no_cpu_baseline:
        res, cores, cn = _cpu_baseline()
        cpu = {"value": round(res[cores][0], 2), "unit": "MLUPS", "cores": cores, "kind": "port",
               "value_1core": round(res[1][0], 2), "host": _host_info(),
               "sample": "%d^3 D3Q19 MRT Taylor-Green, FP64 AoS, C restatement of pylbm's generated Cython "
                         "one_time_step + periodic ghost fill (oracle/cgen.py), %d steps on %d OpenMP threads, "
                         "%d steps on 1 thread" % (cn, res[cores][1], cores, res[1][1])}

    del sim
    if not args.no_configs:
        extra["configs"] = _configs_leg(args.config, args.config_steps)
    if not args.no_surrogate:
        extra["surrogate"] = _surrogate_leg(args.sur_traj, args.sur_steps, args.sur_train, elem=args.sur_elem)

    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "D3Q19 MRT periodic Taylor-Green vortex, %d^3 cells, u0=0.04, nu=0.005, "
                               "FP32 populations, FP32 arithmetic" % n,
                   "cells": cells, "storage": "float32", "device": rt.device_name,
                   "wall_s_timed": round(wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": "fused_step<float,true> (pull + f2m + MRT relax + m2f)",
                     "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
