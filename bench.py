"""Benchmark of the KDK Schroedinger-Poisson step on one MI355X.

Workload (the headline configuration): a 256^3 complex64 psi (seeded
filtered Gaussian, mean |psi|^2 = 1) plus 2^18 equal-mass FP64 stars
(uniform positions, N(0, 0.05^2) velocities, total mass 0.1, rho_bar = 1.1),
dt = 5e-4.  One step = one full kick-drift-kick step of psi and the stars.

Prints ONE JSON line: cell-updates/s of the fused gfx950 path, the roofline
of its dominant kernel (per-stage HIP-event timing), the plain complex128
PyTorch path on the same GPU, accuracy against the FP64 (complex128) path
after the same number of steps, and the CPU restatement of the reference
path (FP64 NumPy/SciPy, timed on a bounded sample on this host).

Usage: python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--n 256]
                       [--stars 262144] [--no-cpu] [--no-plain] [--no-accuracy]
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def affinity_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_threads():
    """Threads for the CPU restatement: every core this process may run on,
    or fewer when the host states a smaller CPU share for the job in
    OMP_NUM_THREADS (a shared GPU host gives each GPU's job its own share)."""
    n = affinity_cpus()
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def stage_bytes(name, nx, ny, nz, c, r, n_p, stars, gather=True, fdep=True, rho_write=False):
    """Algorithmic HBM bytes of one launch of a stage (read + write once).

    Stars use the tile-binned deposit: it overwrites every cell of rho_s
    (N r), so the source pass only reads rho_s; its other bytes are the
    star positions and registrations (about 64 B per star on the deposit's
    own tiles).  With the fused star gather (the default for stars steps)
    the Poisson x pass writes V and, instead of the acceleration grid, each
    star's eight corner accelerations (24 r B), reading its two column-tile
    registrations (index and corner descriptor, 12 B each); the deposit's
    scatter writes those descriptors (about 112 B per star in all).  With
    the fused deposit as well (the default) each star has one registration
    in the tile of its lower y row (two where its z pair crosses a tile
    edge), which the tile and the tile above both read: the source pass
    reads it twice and the star's eight corner weights (24 + 8 r B per
    star), and the Poisson x pass reads it twice as well; the source pass
    writes rho_s (N r) only with rho_write (QJ_OPT_FUSED_DEPOSIT 1), and the
    deposit stage is the scan and the scatter (position, rank, registration,
    descriptor and weights: 40 + 8 r B per star)."""
    N = nx * ny * nz
    Nh = (nx // 2 + 1) * ny * nz
    src = r
    fused = stars and gather and n_p > 0
    fd = fused and fdep
    if fd and not rho_write:
        src = 0  # rho_s neither read nor written
    table = {
        "drift_z_kick": N * (2 * c + r),
        "drift_z": 2 * N * c, "drift_y": 2 * N * c, "drift_x": 2 * N * c,
        "drift_x_source": 2 * N * c + N * src + Nh * c + (n_p * (24 + 8 * r) if fd else 0),
        "source_x": N * c + N * src + Nh * c + (n_p * (24 + 8 * r) if fd else 0),
        "fft_y_half": 2 * Nh * c,
        "poisson_z": 3 * Nh * c,
        "poisson_y": 5 * Nh * c,
        # V plane + (a_x, a_y, a_z) cells, or V + the stars' corner values
        "poisson_x": 3 * Nh * c + (N * r + n_p * (24 * r + 24) if fused else 4 * N * r),
        "deposit": n_p * (40 + 8 * r) if fd else N * r + n_p * (112 if fused else 64),
        "star_step": n_p * (96 + 8 * 3 * r + 8),
        "star_kick": n_p * (72 + 8 * 3 * r),
        "kick": N * (2 * c + r),
    }
    return table.get(name)


def stream_copy_gbs(dev):
    n = 1 << 28  # 1 GiB of float32
    a = torch.empty(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    reps = 20
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    del a, b
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def gpu_psi(shape, seed, dev, sigma_k=16 * np.pi):
    """Seeded filtered-Gaussian psi built on the device (torch generator):
    Re, Im ~ N(0,1), amplitude filter exp(-|k|^2/(2 sigma_k^2)), mean |psi|^2 = 1."""
    g = torch.Generator(device=dev).manual_seed(seed)
    f = torch.complex(torch.randn(shape, generator=g, device=dev, dtype=torch.float64),
                      torch.randn(shape, generator=g, device=dev, dtype=torch.float64))
    k = [2 * np.pi * torch.fft.fftfreq(n, 1.0 / n, device=dev, dtype=torch.float64)
         for n in shape]
    k2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    f = torch.fft.ifftn(torch.fft.fftn(f) * torch.exp(-k2 / (2 * sigma_k ** 2)))
    return (f / f.abs().pow(2).mean().sqrt()).to(torch.complex64)


def ensemble_leg(qj, dev, phys, nsteps=50):
    """16 x 64^3 + 24 x 128^3 + 8 x 256^3 complex64 boxes, 2^14 stars each,
    per-box seed 17+b and dt ~ U[2e-4, 1e-3]; 50 KDK steps."""
    sizes = [64] * 16 + [128] * 24 + [256] * 8
    rng = np.random.default_rng(4242)
    dts = rng.uniform(2e-4, 1e-3, len(sizes))
    n_p = 1 << 14
    log(f"ensemble: {len(sizes)} boxes (16x64^3, 24x128^3, 8x256^3), {n_p} stars each")
    sims, ics = [], []
    for b, (n, dtb) in enumerate(zip(sizes, dts)):
        pl = qj.Plan(n, dtype=torch.complex64, device=dev)
        x, v, m = qj.ic.star_ic(n_p, 1017 + b)
        ic = (gpu_psi((n, n, n), 17 + b, dev), torch.as_tensor(x).to(dev),
              torch.as_tensor(v).to(dev), torch.full((n_p,), m, dtype=torch.float64, device=dev))
        ics.append(ic)
        sims.append(qj.KDKSimulation(pl, phys, float(dtb), ic[0], ic[1], ic[2], mass=m))
    cells = sum(n ** 3 for n in sizes)
    ens = qj.Ensemble(sims, mode="eager")
    ens.step(1)
    torch.cuda.synchronize()
    a, b_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    modes = {}
    for m in ("eager", "batched", "graph", "graph_fork"):
        ens.set_mode(m)
        ens.step(1)
        torch.cuda.synchronize()
        a.record()
        ens.step(10)
        b_.record()
        torch.cuda.synchronize()
        modes[m] = a.elapsed_time(b_) / 10
        log(f"ensemble mode {m}: {modes[m]:.3f} ms/step")
    t0 = time.perf_counter()
    tuned = ens.autotune(reps=3)
    t_tune = time.perf_counter() - t0
    knob_names = [k[0] for k in qj.Ensemble.KNOBS]
    log(f"ensemble autotune: {tuned['mode']} " +
        " ".join(f"{k} {tuned.get(k)}" for k in knob_names) + f" ({t_tune:.1f} s)")
    ens.step(1)
    torch.cuda.synchronize()
    a.record()
    ens.step(nsteps)
    ens.finalize()
    b_.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b_) / nsteps
    groups = [(len(m), nl) for m, nl in ens.groups()]
    ens.close()
    del ens, sims
    torch.cuda.empty_cache()

    # forward + adjoint of the same 48 boxes (checkpoint every 10 steps)
    ad = importlib.import_module("quantum-jax_amd.adjoint")
    seg = 10
    plans = [qj.Plan(n, dtype=torch.complex64, device=dev) for n in sizes]
    tgts = [gpu_psi((n, n, n), 5017 + b, dev).abs().pow(2).to(torch.float32)
            for b, n in enumerate(sizes)]

    def run(grad):
        ps = [ic[0].clone().requires_grad_(grad) for ic in ics]
        ms = [ic[3].clone().requires_grad_(grad) for ic in ics]
        with torch.set_grad_enabled(grad):
            po, xo, _ = ad.ensemble_kdk_run(ps, [ic[1] for ic in ics], [ic[2] for ic in ics], ms,
                                            plans, [phys] * len(sizes), list(dts), nsteps, seg,
                                            mode=tuned["mode"])
            L = sum(ad.density_loss(po[b], xo[b], ms[b], tgts[b], plans[b])
                    for b in range(len(sizes)))
            if grad:
                L.backward()
                ok = all(bool(torch.isfinite(p.grad).all()) for p in ps)
                return ok
        return True

    log(f"ensemble adjoint: {nsteps} steps, checkpoint every {seg}")
    run(False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    run(False)
    torch.cuda.synchronize()
    t_f = time.perf_counter() - t
    run(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ok = run(True)
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t
    log(f"ensemble adjoint: forward {1e3 * t_f / nsteps:.3f} ms/step, forward + adjoint "
        f"{1e3 * t_fb / nsteps:.3f} ms/step")
    adj = {"workload": f"same 48 boxes, {nsteps} steps, checkpoint every {seg}, loss sum over boxes "
                       "of sum((|psi|^2 + rho_s - rho_target)^2), grad w.r.t. psi0 and masses",
           "forward_ms_per_step": 1e3 * t_f / nsteps,
           "forward_plus_adjoint_ms_per_step": 1e3 * t_fb / nsteps,
           "cost_ratio": t_fb / t_f, "grad_finite": ok}
    del plans, tgts, ics
    torch.cuda.empty_cache()
    return {
        "workload": "48 complex64 boxes: 16x64^3 + 24x128^3 + 8x256^3 (188,743,680 cells), "
                    f"2^14 FP64 stars each, per-box seed 17+b, dt ~ U[2e-4, 1e-3], {nsteps} steps",
        "value": cells / (ms * 1e-3), "unit": "cell-updates/s", "ms_per_step": ms,
        "mode_ms_per_step": modes,
        "speedup_vs_eager": modes["eager"] / ms,
        "autotuned": dict({"mode": tuned["mode"]}, **{k: tuned.get(k) for k in knob_names},
                          tuning_s=t_tune),
        "launch_groups": groups,
        "adjoint": adj,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--stars", type=int, default=1 << 18)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-gas", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--side-stream", type=int, default=1, choices=(0, 1),
                    help="star work of the timed steps on the plan's second stream (1) or on "
                         "the caller's stream (0; every kernel then runs alone)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the decomposition of one grid over several "
              "GPUs (x slabs with local y/z sweeps and all-to-all transposes for the x "
              "sweeps) is not built, so --gpus must be 1.",
              file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)

    qj = importlib.import_module("quantum-jax_amd")
    from importlib import import_module
    lib_mod = import_module("quantum-jax_amd._lib")
    diag_mod = import_module("quantum-jax_amd.diagnostics")
    dev = torch.device("cuda", 0)
    n, n_p, dt = args.n, args.stars, args.dt
    N = n ** 3
    phys = qj.Physics(rho_bar=1.1)

    log(f"IC: {n}^3 psi seed 17, {n_p} stars seed 1017")
    psi0 = qj.ic.filtered_gaussian_psi((n, n, n), 17)
    pos0, vel0, m_s = qj.ic.star_ic(n_p, 1017)
    psi_d = torch.as_tensor(psi0).to(dev, torch.complex64)
    pos_d = torch.as_tensor(pos0).to(dev)
    vel_d = torch.as_tensor(vel0).to(dev)

    plan = qj.Plan(n, dtype=torch.complex64, device=dev)
    plan.set_option(lib_mod.OPT_SIDE_STREAM, args.side_stream)
    sim = qj.KDKSimulation(plan, phys, dt, psi_d, pos_d, vel_d, mass=m_s)

    # ---- warmup + timed region ------------------------------------------
    log(f"warmup {args.warmup} steps")
    sim.step(args.warmup)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"timed {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    sim.step(args.steps)
    sim.finalize()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    value = N * args.steps / (ms_total * 1e-3)
    log(f"{ms_step:.4f} ms/step, {value:.4e} cell-updates/s (host wall {wall:.3f}s)")

    # ---- the other x schedules (fused x drift + Poisson source, split, and
    # the automatic choice between them) ---------------------------------
    fuse0 = plan.get_option(lib_mod.OPT_FUSE_X_SOURCE)
    schedules = {f"fuse_x_source={fuse0}": ms_step}
    for alt in (0, 1, 2):
        if alt == fuse0:
            continue
        plan.set_option(lib_mod.OPT_FUSE_X_SOURCE, alt)
        sim.step(3)
        torch.cuda.synchronize()
        ev0.record()
        sim.step(args.steps)
        sim.finalize()
        ev1.record()
        torch.cuda.synchronize()
        ms_alt = ev0.elapsed_time(ev1) / args.steps
        log(f"x schedule fuse={alt}: {ms_alt:.4f} ms/step (fuse={fuse0}: {ms_step:.4f})")
        schedules[f"fuse_x_source={alt}"] = ms_alt
    plan.set_option(lib_mod.OPT_FUSE_X_SOURCE, fuse0)

    # ---- star kernels on the caller's stream instead of the side stream --
    side0 = plan.get_option(lib_mod.OPT_SIDE_STREAM)
    plan.set_option(lib_mod.OPT_SIDE_STREAM, 1 - side0)
    sim.step(3)
    torch.cuda.synchronize()
    ev0.record()
    sim.step(args.steps)
    sim.finalize()
    ev1.record()
    torch.cuda.synchronize()
    ms_side = ev0.elapsed_time(ev1) / args.steps
    log(f"side stream {1 - side0}: {ms_side:.4f} ms/step (side stream {side0}: {ms_step:.4f})")
    plan.set_option(lib_mod.OPT_SIDE_STREAM, side0)
    schedules.update({f"side_stream={side0}": ms_step, f"side_stream={1 - side0}": ms_side})

    # ---- per-stage timing pass (HIP events around every launch) ---------
    # run with every stage on one stream: beside the psi sweeps, the star
    # kernels of the side stream would stretch each other's event intervals
    L = lib_mod.lib()
    nst = L.qj_stage_count()
    names = [L.qj_stage_name(k).decode() for k in range(nst)]
    side0 = plan.get_option(lib_mod.OPT_SIDE_STREAM)
    plan.set_option(lib_mod.OPT_SIDE_STREAM, 0)
    L.qj_timing_enable(plan.handle, 1)
    tsteps = max(10, min(50, args.steps))
    sim.step(tsteps)
    import ctypes
    ms_arr = (ctypes.c_double * nst)()
    cnt_arr = (ctypes.c_longlong * nst)()
    lib_mod.check(L.qj_timing_read(plan.handle, ms_arr, cnt_arr, nst), "timing_read")
    L.qj_timing_enable(plan.handle, 0)
    plan.set_option(lib_mod.OPT_SIDE_STREAM, side0)
    c_b, r_b = 8, 4
    stages = {}
    for k in range(nst):
        if cnt_arr[k]:
            avg = ms_arr[k] / cnt_arr[k]
            b = stage_bytes(names[k], n, n, n, c_b, r_b, n_p, True,
                            bool(plan.get_option(lib_mod.OPT_STAR_GATHER)),
                            bool(plan.get_option(lib_mod.OPT_FUSED_DEPOSIT)) and
                            not plan.get_option(lib_mod.OPT_DETERMINISTIC_DEPOSIT),
                            plan.get_option(lib_mod.OPT_FUSED_DEPOSIT) == 1)
            stages[names[k]] = {
                "launches_per_step": round(cnt_arr[k] / tsteps, 3),
                "avg_ms": round(avg, 5),
                "bytes_per_launch": b,
                "gbs": round(b / (avg * 1e-3) / 1e9, 1) if b else None,
            }
    dom = max(stages, key=lambda s: ms_arr[names.index(s)])
    dstat = stages[dom]
    achieved = dstat["gbs"]
    log(f"dominant kernel {dom}: {dstat['avg_ms']} ms, {achieved} GB/s")
    step_bytes = sum(v["bytes_per_launch"] * v["launches_per_step"] for v in stages.values()
                     if v["bytes_per_launch"])
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    # (scaled by a 1 GiB copy) for this workload, committed under profiles/
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "kdk_stars_256_hbm_traffic.json")
    if n == 256 and n_p == (1 << 18) and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh)["bytes_per_launch"].get(dom)
    out = {
        "metric": "KDK cell-updates/s at 256^3 (c64 vs plain c128, same GPU) + adjoint step cost",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded filtered-Gaussian psi, uniform FP64 stars)",
        "config": {"workload": f"stars {n}^3 c64 psi + {n_p} FP64 stars, dt={dt}",
                   "grid": [n, n, n], "stars": n_p, "dt": dt, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": dstat["bytes_per_launch"]},
        "step_bytes_algorithmic": int(step_bytes),
        "step_gbs_effective": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
        "schedule_s_fraction": round(227.4 * value / 8.0e12, 4) if n == 256 else None,
        "stages": stages,
        "stages_note": "per-launch HIP events with all stages on one stream (side stream off); "
                       "the timed steps run the star kernels on the plan's second stream",
        "schedules_ms_per_step": schedules,
    }
    del sim
    torch.cuda.empty_cache()

    out["stream_copy_gbs"] = round(stream_copy_gbs(dev), 1)

    # ---- accuracy vs the FP64 (complex128) path ----------------------------
    if not args.no_accuracy:
        log("accuracy: c64 and c128 fused paths from the same IC")
        nacc = args.steps
        res = {}
        for tag, cd in (("c64", torch.complex64), ("c128", torch.complex128)):
            pl = qj.Plan(n, dtype=cd, device=dev)
            s = qj.KDKSimulation(pl, phys, dt, torch.as_tensor(psi0).to(dev, cd),
                                 torch.as_tensor(pos0).to(dev), torch.as_tensor(vel0).to(dev),
                                 mass=m_s)
            E0 = diag_mod.energy(pl, phys, s.psi, s.pos, s.vel, m_s)
            M0 = diag_mod.mass(pl, s.psi, torch.full((n_p,), m_s, dtype=torch.float64,
                                                     device=dev))
            s.step(nacc)
            st = s.state()
            E1 = diag_mod.energy(pl, phys, st["psi"], st["pos"], st["vel"], m_s)
            M1 = diag_mod.mass(pl, st["psi"], torch.full((n_p,), m_s, dtype=torch.float64,
                                                         device=dev))
            res[tag] = dict(psi=st["psi"].to(torch.complex128), pos=st["pos"].clone(),
                            E0=E0, E1=E1, M0=M0, M1=M1)
            del s, pl
            torch.cuda.empty_cache()
        a, b = res["c64"], res["c128"]
        err = ((a["psi"] - b["psi"]).norm() / b["psi"].norm()).item()
        dpos = ((a["pos"] - b["pos"] + 0.5) % 1.0 - 0.5).abs().max().item()
        out["accuracy"] = {
            "steps": nacc,
            "psi_rel_l2_vs_fp64": err,
            "star_pos_max_abs_vs_fp64": dpos,
            "mass_drift_c64": abs(a["M1"] - a["M0"]) / a["M0"],
            "mass_drift_c128": abs(b["M1"] - b["M0"]) / b["M0"],
            "energy_drift_c64": abs(a["E1"] - a["E0"]) / abs(a["E0"]),
            "energy_drift_c128": abs(b["E1"] - b["E0"]) / abs(b["E0"]),
            "fp64_reference": "the complex128 build of the same fused kernels, pinned to the "
                              "FP64 NumPy oracle at 1e-11 in tests/test_gpu_parity.py",
        }
        log(f"accuracy: {out['accuracy']}")
        del res, a, b
        torch.cuda.empty_cache()

    # ---- HIP complex128 and the plain complex128 PyTorch path ----------------
    if not args.no_plain:
        k128 = max(5, min(20, args.steps // 10))
        pl = qj.Plan(n, dtype=torch.complex128, device=dev)
        s = qj.KDKSimulation(pl, phys, dt, torch.as_tensor(psi0).to(dev, torch.complex128),
                             torch.as_tensor(pos0).to(dev), torch.as_tensor(vel0).to(dev),
                             mass=m_s)
        s.step(2)
        torch.cuda.synchronize()
        ev0.record()
        s.step(k128)
        s.finalize()
        ev1.record()
        torch.cuda.synchronize()
        hip128 = N * k128 / (ev0.elapsed_time(ev1) * 1e-3)
        del s, pl
        torch.cuda.empty_cache()
        plain = importlib.import_module("quantum-jax_amd.plain")
        pk = plain.PlainKDK((n, n, n), dtype=torch.complex128, device=dev, rho_bar=1.1)
        p = torch.as_tensor(psi0).to(dev, torch.complex128)
        x = torch.as_tensor(pos0).to(dev)
        v = torch.as_tensor(vel0).to(dev)
        amax = torch.zeros((), dtype=torch.float64, device=dev)
        p, x, v, amax, _ = pk.step_stars(p, x, v, m_s, dt, amax)
        kp = 3
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(kp):
            p, x, v, amax, _ = pk.step_stars(p, x, v, m_s, dt, amax)
        ev1.record()
        torch.cuda.synchronize()
        plain128 = N * kp / (ev0.elapsed_time(ev1) * 1e-3)
        out["plain_c128_same_gpu"] = {
            "value": plain128, "unit": "cell-updates/s", "steps": kp,
            "what": "reference operator sequence in unfused torch ops (torch.fft, index_add_)",
        }
        out["hip_c128"] = {"value": hip128, "unit": "cell-updates/s", "steps": k128}
        out["speedup_vs_plain_c128"] = value / plain128
        log(f"plain c128 {plain128:.3e}, hip c128 {hip128:.3e}, speedup {value / plain128:.1f}x")
        del p, x, v, pk
        torch.cuda.empty_cache()

    # ---- quantum + isothermal gas workload: 256^3, 100 steps --------------
    if not args.no_gas:
        ng, kg, dtg, csg = n, 100, dt, 0.1
        log(f"gas: {ng}^3 c64 psi + FP32 gas, {kg} steps")
        rho_g0, vx0, vy0, vz0 = qj.ic.gas_ic((ng, ng, ng), 17)
        rb = float(np.mean(np.abs(psi0) ** 2 + rho_g0))
        # the benchmark's gas scheme: MUSCL-Hancock, minmod, HLL flux, CFL 0.4
        # (dt = min(dt, 0.4 dx / max(cs + |v|)) from the device each step)
        cflg = 0.4
        gphys = qj.Physics(rho_bar=rb, cs=csg, flux="hll")
        res = {}
        for tag, cd in (("c64", torch.complex64), ("c128", torch.complex128)):
            rd = torch.float32 if cd == torch.complex64 else torch.float64
            pg = qj.Plan(ng, dtype=cd, device=dev)
            gas = [torch.as_tensor(a).to(dev, rd) for a in (rho_g0, vx0, vy0, vz0)]
            sg = qj.KDKSimulation(pg, gphys, dtg, torch.as_tensor(psi0).to(dev, cd), gas=gas)
            M0 = diag_mod.mass(pg, sg.psi, rho_gas=sg.gas_fields()[0])
            sg.step_cfl(2, cfl=cflg)
            torch.cuda.synchronize()
            ev0.record()
            sg.step_cfl(kg, cfl=cflg)
            sg.finalize()
            ev1.record()
            torch.cuda.synchronize()
            ms_g = ev0.elapsed_time(ev1) / kg
            dts_g = sg.dts[2:]
            gf = sg.gas_fields()
            vmax = max(float(v.abs().max().item()) for v in gf[1:])
            res[tag] = dict(ms=ms_g, psi=sg.psi.to(torch.complex128).clone(),
                            rho=gf[0].double().clone(), dt_min=min(dts_g), dt_max=max(dts_g),
                            M0=M0, M1=diag_mod.mass(pg, sg.psi, rho_gas=gf[0]), vmax=vmax)
            if tag == "c64":   # stage timing after the state was taken
                L = lib_mod.lib()
                L.qj_timing_enable(pg.handle, 1)
                sg.step(10)
                lib_mod.check(L.qj_timing_read(pg.handle, ms_arr, cnt_arr, nst), "timing_read")
                L.qj_timing_enable(pg.handle, 0)
                gst = {names[k]: round(ms_arr[k] / cnt_arr[k], 5) for k in range(nst) if cnt_arr[k]}
                # the reference's own scheme: Rusanov flux of the averaged
                # state at the fixed dt, no per-step CFL read-back
                sr = qj.KDKSimulation(pg, qj.Physics(rho_bar=rb, cs=csg), dtg,
                                      torch.as_tensor(psi0).to(dev, cd),
                                      gas=[torch.as_tensor(a).to(dev, rd)
                                           for a in (rho_g0, vx0, vy0, vz0)])
                sr.step(2)
                torch.cuda.synchronize()
                ev0.record()
                sr.step(kg)
                sr.finalize()
                ev1.record()
                torch.cuda.synchronize()
                ms_rus = ev0.elapsed_time(ev1) / kg
                del sr
            del sg, pg, gas
            torch.cuda.empty_cache()
        a, b = res["c64"], res["c128"]
        dxg = 1.0 / ng
        out["gas"] = {
            "workload": f"{ng}^3 c64 psi + FP32 isothermal gas rho=0.2(1+0.1 g), v=0, cs={csg}, "
                        f"MUSCL-Hancock + minmod + HLL flux, CFL {cflg} with dt = min({dtg}, "
                        f"{cflg} dx / max(cs + |v|)) from the device each step, {kg} steps "
                        f"(after 2 warm-up steps), rho_bar = IC mean",
            "value": ng ** 3 / (a["ms"] * 1e-3), "unit": "cell-updates/s", "ms_per_step": a["ms"],
            "c128_ms_per_step": b["ms"],
            "dt_range": [a["dt_min"], a["dt_max"]],
            "rusanov_fixed_dt_ms_per_step": ms_rus,
            "stage_ms": gst,
            "psi_rel_l2_vs_fp64": ((a["psi"] - b["psi"]).norm() / b["psi"].norm()).item(),
            "gas_rho_rel_l2_vs_fp64": ((a["rho"] - b["rho"]).norm() / b["rho"].norm()).item(),
            "mass_drift_c64": abs(a["M1"] - a["M0"]) / a["M0"],
            "mass_drift_c128": abs(b["M1"] - b["M0"]) / b["M0"],
            "max_cfl": a["dt_max"] * (csg + a["vmax"]) / dxg,
        }
        log(f"gas: {out['gas']}")
        del res, a, b
        torch.cuda.empty_cache()

    # ---- batched ensemble: 16 x 64^3 + 24 x 128^3 + 8 x 256^3 boxes ------------
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(qj, dev, phys)
        torch.cuda.empty_cache()

    # ---- adjoint step cost: 128^3 c64 + 2^16 stars, 256 steps, checkpoint 16 --
    if not args.no_adjoint:
        ad = importlib.import_module("quantum-jax_amd.adjoint")
        na, npa, nsa, seg = 128, 1 << 16, 256, 16
        log(f"adjoint: {na}^3 c64 + {npa} stars, {nsa} steps, checkpoint every {seg}")
        pa = qj.Plan(na, dtype=torch.complex64, device=dev)
        psa = torch.as_tensor(qj.ic.filtered_gaussian_psi((na, na, na), 17)).to(dev,
                                                                                 torch.complex64)
        x0, v0, msa = qj.ic.star_ic(npa, 1017)
        xa = torch.as_tensor(x0).to(dev)
        va = torch.as_tensor(v0).to(dev)
        ma = torch.full((npa,), msa, dtype=torch.float64, device=dev)
        tgt = torch.as_tensor(qj.ic.filtered_gaussian_psi((na, na, na), 2017)).abs().pow(2).to(
            dev, torch.float32)

        def fwd_only():
            with torch.no_grad():
                out = ad.kdk_run(psa, xa, va, ma, pa, phys, dt, nsa, seg)
                return ad.density_loss(out[0], out[1], ma, tgt, pa)

        def fwd_bwd():
            p_ = psa.clone().requires_grad_(True)
            m_ = ma.clone().requires_grad_(True)
            out = ad.kdk_run(p_, xa, va, m_, pa, phys, dt, nsa, seg)
            L_ = ad.density_loss(out[0], out[1], m_, tgt, pa)
            L_.backward()
            return p_.grad, m_.grad

        fwd_only()
        fwd_bwd()
        torch.cuda.synchronize()
        t = time.perf_counter()
        fwd_only()
        torch.cuda.synchronize()
        t_f = time.perf_counter() - t
        t = time.perf_counter()
        gp, gm = fwd_bwd()
        torch.cuda.synchronize()
        t_fb = time.perf_counter() - t
        out["adjoint"] = {
            "workload": f"{na}^3 c64 + {npa} stars, {nsa} steps, dt={dt}, checkpoint every {seg}, "
                        "loss sum((|psi|^2 + rho_s - rho_target)^2), grad w.r.t. psi0 and masses",
            "forward_ms_per_step": 1e3 * t_f / nsa,
            "forward_plus_adjoint_ms_per_step": 1e3 * t_fb / nsa,
            "cost_ratio": t_fb / t_f,
            "grad_finite": bool(torch.isfinite(gp).all().item() and torch.isfinite(gm).all().item()),
        }
        log(f"adjoint: {out['adjoint']}")
        del pa, psa, gp, gm
        torch.cuda.empty_cache()

    # ---- CPU restatement of the reference path (bounded sample) -------------
    if not args.no_cpu:
        import oracle as O
        thr = cpu_threads()
        O.set_workers(thr)
        g = O.Grid(n, n, n)
        ph = O.Physics(rho_bar=1.1)
        st = (psi0, pos0, vel0, 0.0, 0.0, np.inf)
        log(f"cpu baseline: oracle step at {n}^3, {thr} threads")
        st = O.compute_step_stars(*st, dt, m_s, g, ph)   # warm-up
        times = []
        for _ in range(3):
            t = time.perf_counter()
            st = O.compute_step_stars(*st, dt, m_s, g, ph)
            times.append(time.perf_counter() - t)
        tmed = float(np.median(times))
        # the same step with one FFT worker (the scalar figure)
        O.set_workers(1)
        t = time.perf_counter()
        st = O.compute_step_stars(*st, dt, m_s, g, ph)
        t1 = time.perf_counter() - t
        O.set_workers(thr)
        n_or = 2 + len(times)
        out["cpu_baseline"] = {
            "value": N / tmed, "unit": "cell-updates/s", "cores": thr, "kind": "port",
            "sample": f"1 warm-up + {len(times)} timed steps of the same {n}^3 + {n_p}-star "
                      f"workload, FP64 NumPy/SciPy restatement of the reference operator order "
                      f"(unmerged kicks, 14 FFTs/step, scipy.fft workers={thr}); "
                      f"median {tmed:.2f} s/step; then 1 step with workers=1: {t1:.2f} s; "
                      f"CPU: {cpu_model()}",
            "value_workers_1": N / t1,
            "affinity_cpus": affinity_cpus(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        }
        log(f"cpu: {tmed:.2f} s/step ({thr} workers), {t1:.2f} s/step (1 worker)")
        # the oracle's state after those steps checks the HIP builds stepped
        # from the same IC at this size
        ref_psi = torch.as_tensor(st[0])
        ref_pos = st[1]
        vs = {"steps": n_or}
        for tag, cd in (("c64", torch.complex64), ("c128", torch.complex128)):
            pl = qj.Plan(n, dtype=cd, device=dev)
            s = qj.KDKSimulation(pl, phys, dt, torch.as_tensor(psi0).to(dev, cd),
                                 torch.as_tensor(pos0).to(dev), torch.as_tensor(vel0).to(dev),
                                 mass=m_s)
            s.step(n_or)
            sst = s.state()
            p64 = sst["psi"].to(torch.complex128).cpu()
            vs[f"psi_rel_l2_{tag}"] = ((p64 - ref_psi).norm() / ref_psi.norm()).item()
            dp = (sst["pos"].cpu().numpy() - ref_pos + 0.5) % 1.0 - 0.5
            vs[f"star_pos_max_abs_{tag}"] = float(np.abs(dp).max())
            del s, pl, sst
            torch.cuda.empty_cache()
        out["accuracy_vs_oracle"] = vs
        log(f"vs the FP64 oracle at {n}^3: {vs}")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
