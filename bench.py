"""Benchmark of the per-timestep hot loop on one MI355X.

Workload: 2^24 particles on a 64^3 x 16 x 16 grid, box L = 50, dt = 1e-3,
D_t = 0.1, D_r = 0.5, alpha = -1 (pushers), uniform positions and isotropic
orientations with a 5% sin(2 pi x / L) polar perturbation, Philox noise from
seed 2.  `value` is the FP32 mixed-precision path (compensated positions, FP32
FFT); the plain FP64 GPU path of the same physics is timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (1 GPU, 2^24 particles, 64³×16×16) + %HBM peak; FP32 vs FP64"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step of the update kernel (FP32 + in-kernel
# Philox: read x,y,z,phi,theta (20 B) + 3 compensation words of 16 bits (6 B),
# write the same) and of the FP64 plain update (read 40 + write 40; noise is
# generated in-kernel).  The byte accounting of the path's survey assumed
# 32-bit compensation words, 64 B per FP32 particle-step; the fraction of the
# HBM peak at that figure is reported beside (frac_at_64B).
UPDATE_BYTES = {"f32": 52.0, "f64": 80.0}
SURVEY_BYTES_F32 = 64.0
# The plain path's bin stage reads the five FP64 coordinates (40 B per
# particle) and writes the count grid (4 B per bin) and the moments (16 B per
# cell and component); on the tiled path the binning of the new state is fused
# into the update.
BIN_BYTES = {"f64": 40.0}


def bin_stage_bytes(n, grid=(64, 64, 64, 16, 16), nm=6):
    """Algorithmic bytes of one plain-path bin stage."""
    ns, na = grid[0] * grid[1] * grid[2], grid[3] * grid[4]
    return BIN_BYTES["f64"] * n + 4.0 * ns * na + 16.0 * nm * ns


def hbm_frac(nbytes, n, ms):
    """Fraction of the HBM peak that nbytes per particle at n particles in ms is."""
    return nbytes * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS if ms and ms > 0 else None


def make(prec, n, binning=None, **extra):
    import stochasticsampling_amd as ss

    kw = dict(extra)
    if binning is not None:
        kw["binning"] = binning
    return ss.make_params(
        box=(50.0, 50.0, 50.0), grid=(64, 64, 64, 16, 16), n=n, timestep=1e-3, seed=2,
        diffusion_translational=0.1, diffusion_rotational=0.5, stress_active=-1.0, precision=prec, **kw
    )


def time_gpu(prec, parts, steps, warmup, profile_steps, tune=False, **extra):
    """Times exactly `steps` steps after `warmup` untimed ones.  On the tiled
    path the timed window starts with a re-sorting update (ss_resort), so it
    holds ceil(steps / R) re-sorts for a re-sort interval R: never fewer than
    its share of a whole period."""
    import stochasticsampling_amd as ss
    from stochasticsampling_amd import autotune

    p = make(prec, parts.shape[0], **extra)
    if tune:  # tile shape and re-sort interval of the size class (shipped table)
        p = autotune.tuned(p)
    sim = ss.Simulation(p)
    sim.init(parts)
    if warmup:
        sim.step(warmup)
    sim.resort()
    # set-up of the timed run: capture the graphs it replays (runs nothing)
    sim.prepare_steps(steps)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    t1 = time.perf_counter()
    # per-stage HIP-event timing on the simulation's stream, right after the
    # timed steps on the same state
    stages = sim.profile_steps(profile_steps) if profile_steps else {}
    stages["plan"] = sim.get_info()
    parts_out = sim.get_particles_head(4)
    assert np.all(np.isfinite(parts_out))
    sim.close()
    ms = (t1 - t0) * 1e3 / steps
    return ms, stages


def state_error_study(parts, steps):
    """The FP32 tiled step against the FP64 plain step (the GPU path the
    parity tests tie to the FP64 oracle within 1e-10) after ``steps`` steps
    from the same initial condition with the same Philox noise (the same
    52-bit uniforms in both precisions, FP32 rounding them to binary32):
    particle positions (periodic distance, in
    units of the box length), orientations (|n32 - n64| of the unit
    vectors), and u, grad u of the last step's fields (max |a - b| / max |b|).
    Differences in the binning near cell edges are part of what is
    measured."""
    import stochasticsampling_amd as ss
    from stochasticsampling_amd import autotune

    n = parts.shape[0]
    p32 = autotune.tuned(make(ss.FP32, n))
    sims = [ss.Simulation(p32), ss.Simulation(make(ss.FP64, n, binning=ss.BIN_PLAIN_TILED))]
    out = []
    for sim in sims:
        sim.init(parts)
        sim.step(steps)
        u, g = sim.get_flow_field(with_gradient=True)
        out.append((sim.get_particles(), u, g))
        sim.close()
    (a, ua, ga), (b, ub, gb) = out
    L = np.array([p32.box_size.x, p32.box_size.y, p32.box_size.z])
    d = np.abs(a[:, :3] - b[:, :3])
    d = np.minimum(d, L - d) / L

    def unit(p):
        st = np.sin(p[:, 4])
        return np.stack([st * np.cos(p[:, 3]), st * np.sin(p[:, 3]), np.cos(p[:, 4])], 1)

    dn = np.linalg.norm(unit(a) - unit(b), axis=1)
    dp = d.max(axis=1)
    q = [50, 99, 99.99]
    return {"steps": steps, "reference": "FP64 plain GPU path",
            "position_over_L": {"max": float(dp.max()), "rms": float(np.sqrt((d * d).mean())),
                                "percentiles": dict(zip(map(str, q), map(float, np.percentile(dp, q))))},
            "orientation": {"max": float(dn.max()), "rms": float(np.sqrt((dn * dn).mean())),
                            "percentiles": dict(zip(map(str, q), map(float, np.percentile(dn, q))))},
            "u_rel_err": float(np.abs(ua - ub).max() / np.abs(ub).max()),
            "grad_u_rel_err": float(np.abs(ga - gb).max() / np.abs(gb).max()),
            "note": "trajectory divergence after the given steps: both runs draw the same increments "
                    "(to FP32 rounding), and their differences are arithmetic, amplified where the two runs "
                    "bin a particle near a cell edge differently; the arithmetic error of one solve on the "
                    "same distribution is fft_variants.*_rel_err"}


def fft_variant_study(parts, steps):
    """The spectral-stage variants of the FP32 step on the same workload:
    fused radix kernels (the default), radix axis passes, and dense DFT
    products on the matrix cores (three-part split bf16, six products) in
    axis passes.  Timing:
    mean ms per step of each stage over `steps` steps (HIP events).  Accuracy:
    u and grad u of the spectral solve (SpectralSolver::mean_flow_field) in
    FP32 with the variant's transforms against the FP64 solve, both from the
    same distribution (the initial condition binned on the GPU in FP64),
    max |a - b| / max |b|."""
    import stochasticsampling_amd as ss

    n = parts.shape[0]
    ref = ss.Simulation(make(ss.FP64, n))
    ref.init(parts)
    dist = ref.get_distribution()
    ref.close()
    p64 = make(ss.FP64, n)
    u64, g64 = ss.mean_flow_field(p64, dist)
    out = {}
    for name, kw in [("radix_fused", {}), ("radix_axis_passes", {"flags": ss.FLAG_AXIS_PASSES}),
                     ("dft_mfma_axis_passes", {"fft_variant": ss.FFT_DFT})]:
        p = make(ss.FP32, n, **kw)
        u, g = ss.mean_flow_field(p, dist)
        eu = float(np.max(np.abs(u - u64)) / np.max(np.abs(u64)))
        eg = float(np.max(np.abs(g - g64)) / np.max(np.abs(g64)))
        sim = ss.Simulation(p)
        sim.init(parts)
        sim.step(1)
        stages = sim.profile_steps(steps)
        sim.close()
        spectral = sum(v for k, v in stages.items() if k not in ("noise", "update", "sort"))
        out[name] = {"ms_per_step": sum(stages.values()), "spectral_ms": spectral, "stages_ms": stages,
                     "u_rel_err_vs_fp64": eu, "grad_u_rel_err_vs_fp64": eg}
    del dist
    ok = {k: v for k, v in out.items() if v["u_rel_err_vs_fp64"] <= 1e-5 and v["grad_u_rel_err_vs_fp64"] <= 1e-5}
    out["kept"] = min(ok, key=lambda k: ok[k]["spectral_ms"]) if ok else None
    out["gate"] = ("u and grad u of the FP32 solve within 1e-5 of the FP64 solve on the same distribution; "
                   "the fastest spectral stage among the variants passing is kept")
    return out


def host_cores():
    """Cores this process may use: the CPU affinity mask, capped by
    OMP_NUM_THREADS when the environment sets it (the share of the host's
    cores given to one GPU)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(aff, omp) if omp > 0 else aff), {"nproc": os.cpu_count(), "affinity": aff,
                                                   "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def small_case(steps, threads):
    """The CPU-runnable case, FP64: L = 10, 16^3 x 8 x 8, 2^16 particles,
    dt = 1e-3, D_t = 0.1, D_r = 0.5, alpha = -1, seed 1; the GPU (plain FP64
    path) and the CPU oracle (reference-faithful step structure, `threads`
    OpenMP threads) each timed over all `steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    import stochasticsampling_amd as ss

    n = 1 << 16
    p = ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=n, timestep=1e-3, seed=1,
                       diffusion_translational=0.1, diffusion_rotational=0.5, stress_active=-1.0)
    parts = ss.initial_condition(p, amplitude=0.05, seed=1)
    sim = ss.Simulation(p)
    sim.init(parts)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    gpu_s = time.perf_counter() - t0
    sim.close()
    o = Oracle()
    o.set_threads(threads)
    po = ss.make_params(box=(10.0, 10.0, 10.0), grid=(16, 16, 16, 8, 8), n=n, timestep=1e-3, seed=1,
                        diffusion_translational=0.1, diffusion_rotational=0.5, stress_active=-1.0,
                        flags=ss.FLAG_MAGNETIC_ALWAYS)
    osim = o.simulation(po)
    osim.init(parts)
    t0 = time.perf_counter()
    osim.step(steps)
    cpu_s = time.perf_counter() - t0
    return {"workload": "FP64, L=10, 16^3x8x8, 2^16 particles, dt=1e-3, D_t=0.1, D_r=0.5, alpha=-1, seed 1",
            "steps": steps, "gpu": {"value": n * steps / gpu_s, "ms_per_step": gpu_s * 1e3 / steps,
                                    "path": "FP64 plain (tile-ordered count rows)"},
            "cpu": {"value": n * steps / cpu_s, "ms_per_step": cpu_s * 1e3 / steps, "cores": threads,
                    "kind": "port"},
            "gpu_over_cpu": cpu_s / gpu_s, "unit": "particle-steps/s"}


def magnetic_case(parts, steps, warmup, profile_steps):
    """The magnetic suspension at full size: the 2^24-particle, 64^3 x 16 x 16
    workload in FP32 plus an external field of strength 2 along z acting as
    a torque, the dipole-dipole mean field (mu_dd = 0.5) and its force (drag
    0.1), Philox noise from seed 3; FP64 global diagnostics of the final
    state."""
    import stochasticsampling_amd as ss
    from stochasticsampling_amd import autotune

    n = parts.shape[0]
    kw = dict(magnetic_reorientation=2.0, field_direction=(0.0, 0.0, 1.0), magnetic_dipole_dipole=0.5,
              magnetic_drag=0.1)
    p = autotune.tuned(make(ss.FP32, n, **kw))
    p.seed = 3
    sim = ss.Simulation(p)
    sim.init(parts)
    sim.step(warmup)
    sim.resort()
    sim.prepare_steps(steps)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    stages = sim.profile_steps(profile_steps) if profile_steps else {}
    stages["plan"] = sim.get_info()
    diag = sim.get_diagnostics()
    sim.close()
    upd = stages.get("update")
    achieved = UPDATE_BYTES["f32"] * n / (upd * 1e-3) / 1e9 if upd else None
    return {"value": n / (ms * 1e-3), "unit": "particle-steps/s", "ms_per_step": ms, "steps": steps,
            "dtype": "f32", "stages_ms": stages,
            "workload": "2^24 particles, 64^3x16x16, L=50, FP32, field strength 2 along z (torque), "
                        "mu_dd=0.5, drag=0.1, Philox seed 3",
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS if achieved else None,
                         "bytes_per_particle": UPDATE_BYTES["f32"], "launch_ms": upd},
            "mean_polarisation": [float(v) for v in diag["mean_orientation"]]}


def cpu_baseline(parts, threads, steps):
    """The FP64 CPU oracle in reference-faithful mode (serial binning and stress
    contraction, magnetic solve every step, 36 c2c FFTs, parallel update) on
    the same 2^24-particle, 64^3 x 16 x 16 workload: 1 warm-up step, then
    `steps` timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    import stochasticsampling_amd as ss

    o = Oracle()
    o.set_threads(threads)
    p = make(ss.FP64, parts.shape[0])
    p.flags = ss.FLAG_MAGNETIC_ALWAYS
    sim = o.simulation(p)
    sim.init(parts)
    sim.step(1)
    t0 = time.perf_counter()
    sim.step(steps)
    t1 = time.perf_counter()
    per_step = (t1 - t0) / steps
    return parts.shape[0] / per_step, per_step


def ensemble_configs():
    """48 independent systems: alpha uniformly spaced in [-2, 2], D_r cycling
    over {0.1, 0.5, 1.0}, grids cycling over 32^3x8x8, 48^3x12x12, 64^3x16x16
    with box length 0.78 x grid points, particle counts log-uniform in
    [2^18, 2^22] (numpy seed 3), seeds 100-147."""
    import stochasticsampling_amd as ss

    rng = np.random.default_rng(3)
    grids = [(32, 32, 32, 8, 8), (48, 48, 48, 12, 12), (64, 64, 64, 16, 16)]
    alphas = np.linspace(-2.0, 2.0, 48)
    out = []
    for j in range(48):
        g = grids[j % 3]
        L = 0.78 * g[0]
        n = int(round(2.0 ** rng.uniform(18.0, 22.0)))
        out.append(ss.make_params(box=(L, L, L), grid=g, n=n, timestep=1e-3, seed=100 + j,
                                  diffusion_translational=0.1, diffusion_rotational=(0.1, 0.5, 1.0)[j % 3],
                                  stress_active=float(alphas[j]), precision=ss.FP32, binning=ss.BIN_SORTED))
    return out


def time_ensemble(steps, warmup):
    import ctypes

    import stochasticsampling_amd as ss
    from stochasticsampling_amd import autotune

    # tile shapes per grid class and the shared re-sort interval from the
    # ensemble's entry of the tuning table (tuned on the whole ensemble)
    plist = autotune.tuned_ensemble(ensemble_configs())
    for p in plist:
        if not any(p.tile):
            p.tile = (ctypes.c_int32 * 3)(4, 4, 4)
    ens = ss.Ensemble(plist)
    total = 0
    for p, m in zip(plist, ens.members):
        m.init(ss.initial_condition(p, amplitude=0.05, seed=p.seed))
        total += p.number_of_particles
    ens.step(warmup)
    t0 = time.perf_counter()
    ens.step(steps)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    # per-stage HIP events over eager steps: the persistent update launch
    # moves 64 algorithmic bytes per particle of every member
    stages = ens.profile_steps(32)
    ens.close()
    upd = stages.get("update")
    achieved = UPDATE_BYTES["f32"] * total / (upd * 1e-3) / 1e9 if upd else None
    return {"value": total / (ms * 1e-3), "unit": "particle-steps/s", "ms_per_step": ms, "steps": steps,
            "stages_ms": stages,
            "roofline": {"kernel": "ss::k_tile_update_batched", "bound": "hbm", "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS if achieved else None,
                         "bytes_per_particle": UPDATE_BYTES["f32"], "particles": total, "launch_ms": upd},
            "systems": len(plist), "particles": total, "dtype": "f32",
            "workload": "48 systems, alpha in [-2,2], D_r in {0.1,0.5,1}, grids 32^3x8x8/48^3x12x12/"
                        "64^3x16x16, L = 0.78 x grid points, N log-uniform in [2^18, 2^22], seeds 100-147, "
                        "spectral stages batched per grid class, one persistent update launch per step, "
                        "periods replayed from hipGraphs, tile shapes tuned on the ensemble",
            "tiles": {autotune.grid_class(p): list(p.tile) for p in plist},
            "resort_interval": plist[0].resort_interval}


def copy_probe(gib=1.0, reps=10):
    """Device-to-device copy of a known size (torch), for the bandwidth
    reference and the calibration of the HBM byte counters."""
    import torch

    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"probe": "copy", "bytes_per_copy": 2 * 4 * n, "GBps": 2 * 4 * n / dt / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=1 << 24)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="only run a 1 GiB device copy")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-fft-study", action="store_true")
    ap.add_argument("--state-error-steps", type=int, default=1000)
    ap.add_argument("--fft-study-steps", type=int, default=32)
    ap.add_argument("--no-refnoise", action="store_true")
    ap.add_argument("--refnoise-steps", type=int, default=200)
    ap.add_argument("--refnoise-streams", type=int, default=64)
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--small-steps", type=int, default=100)
    ap.add_argument("--no-magnetic", action="store_true")
    ap.add_argument("--magnetic-steps", type=int, default=100)
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU sharding "
              "of this path is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    import stochasticsampling_amd as ss

    if ss.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    n = args.particles
    p64 = make(ss.FP64, n)
    print(f"[bench] initial condition for {n} particles", file=sys.stderr, flush=True)
    parts = ss.initial_condition(p64, amplitude=0.05, seed=2)

    print("[bench] FP32 sorted path", file=sys.stderr, flush=True)
    ms32, st32 = time_gpu(ss.FP32, parts, args.steps, args.warmup, args.profile_steps, tune=True)
    fp64 = None
    if not args.no_fp64:
        print("[bench] FP64 plain path", file=sys.stderr, flush=True)
        # profiled over one re-sort period (16 steps), so the bin stage holds
        # its share of the re-sort
        ms64, st64 = time_gpu(ss.FP64, parts, max(1, args.steps // 4), min(args.warmup, 5),
                              16 if args.profile_steps else 0, binning=ss.BIN_PLAIN_TILED)
        bb = bin_stage_bytes(n)
        fp64 = {"value": n / (ms64 * 1e-3), "ms_per_step": ms64, "steps": max(1, args.steps // 4),
                "dtype": "f64", "binning": "plain_tiled", "stages_ms": st64,
                # the bin stage: per-tile count rows, drifted particles,
                # contraction (and every 16th step the re-sort)
                "hbm_frac": {"bin_stage": bb / (st64["bin"] * 1e-3) / 1e9 / HBM_PEAK_GBS if st64.get("bin") else None,
                             "update": hbm_frac(UPDATE_BYTES["f64"], n, st64.get("update")),
                             "bytes": {"bin_stage": bb, "update_per_particle": UPDATE_BYTES["f64"]}}}
        # the same plain step with the count grid built by global atomics
        print("[bench] FP64 plain path, atomic histogram", file=sys.stderr, flush=True)
        ms64h, st64h = time_gpu(ss.FP64, parts, max(1, args.steps // 4), min(args.warmup, 5),
                                min(args.profile_steps, 10), binning=ss.BIN_HISTOGRAM)
        fp64["histogram"] = {"value": n / (ms64h * 1e-3), "ms_per_step": ms64h, "binning": "histogram",
                             "stages_ms": st64h,
                             "hbm_frac": {"bin_stage": bb / (st64h["bin"] * 1e-3) / 1e9 / HBM_PEAK_GBS
                                          if st64h.get("bin") else None}}
        # the same FP64 step on the tiled path (binning fused into the update)
        print("[bench] FP64 tiled path", file=sys.stderr, flush=True)
        ms64t, st64t = time_gpu(ss.FP64, parts, max(1, args.steps // 4), min(args.warmup, 5),
                                min(args.profile_steps, 10), binning=ss.BIN_SORTED)
        fp64["tiled"] = {"value": n / (ms64t * 1e-3), "ms_per_step": ms64t, "binning": "sorted",
                         "stages_ms": st64t,
                         "hbm_frac": {"update_with_binning": hbm_frac(UPDATE_BYTES["f64"], n, st64t.get("update"))}}

    fftv = None
    if not args.no_fft_study:
        print("[bench] FFT variants", file=sys.stderr, flush=True)
        fftv = fft_variant_study(parts, args.fft_study_steps)

    serr = None
    if args.state_error_steps > 0 and not args.no_fp64:
        print(f"[bench] FP32 against FP64 after {args.state_error_steps} steps", file=sys.stderr, flush=True)
        serr = state_error_study(parts, args.state_error_steps)

    refn = None
    if not args.no_refnoise:
        # the reference's own noise stream (Pcg32 per rayon thread, ziggurat
        # normals) decoded on the GPU, instead of in-kernel Philox
        print(f"[bench] FP32 with the reference's noise stream ({args.refnoise_streams} streams)",
              file=sys.stderr, flush=True)
        msr, str_ = time_gpu(ss.FP32, parts, args.refnoise_steps, min(args.warmup, 10), min(args.profile_steps, 20),
                             tune=True, noise_mode=ss.NOISE_REFERENCE, noise_streams=args.refnoise_streams)
        refn = {"value": n / (msr * 1e-3), "ms_per_step": msr, "steps": args.refnoise_steps,
                "streams": args.refnoise_streams, "noise_ms": str_.get("noise"), "stages_ms": str_}

    ens = None
    if not args.no_ensemble:
        print("[bench] batched ensemble of 48 systems", file=sys.stderr, flush=True)
        ens = time_ensemble(args.ensemble_steps, 16)

    # roofline of the dominant kernel (the particle update); traffic comes from
    # the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this kernel
    upd_ms = st32.get("update", float("nan"))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_tile_update_fp32_2e24_64c16a.json")
    if os.path.exists(tpath) and n == 1 << 24:
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")
    achieved = UPDATE_BYTES["f32"] * n / (upd_ms * 1e-3) / 1e9
    roofline = {
        "kernel": "ss::k_tile_update<float, 6, true, 0, false, false, false>",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "bytes_per_particle": UPDATE_BYTES["f32"],
        "launch_ms": upd_ms,
        "frac_at_64B": SURVEY_BYTES_F32 * n / (upd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }

    mag = None
    if not args.no_magnetic:
        print("[bench] FP32 magnetic suspension", file=sys.stderr, flush=True)
        mag = magnetic_case(parts, args.magnetic_steps, min(args.warmup, 10), min(args.profile_steps, 20))

    cores, core_info = host_cores()
    threads = args.cpu_threads or cores
    small = None
    if not args.no_small:
        print(f"[bench] small case, GPU and CPU oracle ({threads} threads), {args.small_steps} steps",
              file=sys.stderr, flush=True)
        small = small_case(args.small_steps, threads)

    cpu = None
    if not args.no_cpu:
        print(f"[bench] CPU oracle baseline on {threads} threads", file=sys.stderr, flush=True)
        v, per = cpu_baseline(parts, threads, args.cpu_steps)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "host": core_info,
               "sample": f"FP64 oracle, reference-faithful step structure, full workload "
                         f"(2^24 particles, 64^3x16x16), 1 warm-up + {args.cpu_steps} timed steps, "
                         f"{per:.2f} s/step, {threads} OpenMP threads (all cores of this process's share)"}

    value = n / (ms32 * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "2^24 particles, grid 64^3 x 16 x 16, L=50, dt=1e-3, D_t=0.1, D_r=0.5, "
                        "alpha=-1, FP32 states with compensated positions, FP32 FFT, Philox noise "
                        "seed 2, particles grouped by tiles (tile shape and re-sort interval from "
                        "the autotuner's table, see stages_ms.plan); the timed steps start with a "
                        "re-sorting update",
            "particles": n,
            "grid": [64, 64, 64, 16, 16],
            "box": [50.0, 50.0, 50.0],
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_plain": fp64,
        "ensemble": ens,
        "fft_variants": fftv,
        "fp32_vs_fp64": serr,
        "reference_noise": refn,
        "small_case": small,
        "magnetic": mag,
        "stages_ms": st32,
    }
    if fp64:
        line["fp32_over_fp64"] = value / fp64["value"]
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
