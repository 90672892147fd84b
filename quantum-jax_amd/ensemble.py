"""Ensembles of independent boxes stepped together, and their autotuner.

The reference runs one box per process (quantum-stars.py:454 loops
``jax.lax.fori_loop`` over a single state).  An ``Ensemble`` holds many
``KDKSimulation`` members (any mix of grid sizes, precisions, step kinds and
dt) and steps them together through the native ensemble runner: each step
is recorded once per member, members of the same shape are batched into one
launch per stage (blockIdx.z = member), and the step is replayed from a
hipGraph.  Results are bitwise identical to stepping each member alone.

``autotune`` times the schedule candidates on the live members (execution
mode, and per shape group the x-sweep schedule, the Poisson sweep form,
the z sweeps' workgroups and, with stars, the deposit form, the fused star
gather, its column tiles' rows and the fused deposit) and keeps the
fastest;
member states are saved before and restored after, so tuning does not
advance the simulation.
"""

from __future__ import annotations

import ctypes as C
import json
import os

import torch

from . import _lib
from .simulation import KDKSimulation

MODES = {"eager": _lib.ENS_EAGER, "batched": _lib.ENS_BATCHED, "graph": _lib.ENS_GRAPH,
         "graph_fork": _lib.ENS_GRAPH_FORK}

# tuned schedules of this process, keyed by the ensemble's shape signature
_TUNED: dict = {}


def _kind(sim: KDKSimulation) -> str:
    return "stars" if sim.np else ("gas" if sim.gas is not None else "plain")


def shape_key(sim: KDKSimulation) -> tuple:
    """Members with equal keys have launch sequences of the same shape."""
    p = sim.plan
    return (p.shape, str(p.complex_dtype).split(".")[-1], _kind(sim), sim.np,
            sim.mass_t is not None, sim.brho is not None)


class Ensemble:
    """Independent boxes stepped together.

    sims : KDKSimulation members; each needs its own Plan (plans hold
           reduction scratch and phase tables, and batched members run
           concurrently).  dt and the physics of each member are read once,
           here.
    mode : "graph" (default), "graph_fork", "batched" or "eager".
    """

    def __init__(self, sims, mode: str = "graph"):
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("an ensemble needs at least one member")
        plans = [s.plan for s in self.sims]
        if len({id(p) for p in plans}) != len(plans):
            raise ValueError("every ensemble member needs its own Plan")
        dev = plans[0].device
        if any(p.device != dev for p in plans):
            raise ValueError("all members must live on one device")
        self.device = dev
        n = len(self.sims)
        self._plans = (_lib.vp * n)(*[p.handle.value for p in plans])
        self._states = (C.POINTER(_lib.QjState) * n)(*[C.pointer(s._st) for s in self.sims])
        self._params = (_lib.QjParams * n)(*[s._params for s in self.sims])
        self._works = (_lib.vp * n)(*[p.workspace().data_ptr() for p in plans])
        self._h = C.c_void_p()
        _lib.check(_lib.lib().qj_ensemble_create(C.byref(self._h), n, self._plans, self._states,
                                                 self._params, self._works),
                   "qj_ensemble_create")
        self.mode = None
        self.set_mode(mode)

    # -- schedule ------------------------------------------------------------
    def set_mode(self, mode: str):
        if mode not in MODES:
            raise ValueError(f"mode must be one of {sorted(MODES)}")
        _lib.check(_lib.lib().qj_ensemble_set_mode(self._h, MODES[mode]), "qj_ensemble_set_mode")
        self.mode = mode

    def shape_groups(self) -> dict:
        """shape key -> member indices."""
        g: dict = {}
        for i, s in enumerate(self.sims):
            g.setdefault(shape_key(s), []).append(i)
        return g

    def groups(self):
        """Launch groups of the last recorded step: list of (members, launches per step)."""
        L = _lib.lib()
        out = []
        for gi in range(L.qj_ensemble_group_count(self._h)):
            cap = len(self.sims)
            mem = (C.c_int * cap)()
            nm, nl = C.c_int(), C.c_int()
            _lib.check(L.qj_ensemble_group_info(self._h, gi, mem, cap, C.byref(nm), C.byref(nl)),
                       "qj_ensemble_group_info")
            out.append((list(mem[:nm.value]), nl.value))
        return out

    # -- stepping ------------------------------------------------------------
    def step(self, n: int = 1):
        """Advance every member n steps (asynchronous on the current stream)."""
        if n <= 0:
            return self
        with torch.cuda.device(self.device):
            stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            _lib.check(_lib.lib().qj_ensemble_steps(self._h, int(n), stream), "qj_ensemble_steps")
        for s in self.sims:
            s.t += n * s.dt
        return self

    def finalize(self):
        with torch.cuda.device(self.device):
            stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            _lib.check(_lib.lib().qj_ensemble_finalize(self._h, stream), "qj_ensemble_finalize")
        return self

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            torch.cuda.synchronize(self.device)
            _lib.lib().qj_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- autotuning ------------------------------------------------------------
    def signature(self) -> str:
        keys = sorted((str(k), len(v)) for k, v in self.shape_groups().items())
        return json.dumps(keys)

    def _snapshot(self):
        snap = []
        for s in self.sims:
            # star_scratch: after stars steps with the fused star gather the
            # stars' corner accelerations live there instead of in grid
            t = {k: getattr(s, k) for k in ("_psi", "pos", "vel", "_rho_s", "grid", "diag", "brho",
                                            "star_scratch")}
            t = {k: (v.clone() if v is not None else None) for k, v in t.items()}
            if s.gas is not None:
                t["gas"] = [g.clone() for g in s.gas]
                t["gas_tmp"] = [g.clone() for g in s._gas_tmp]
            snap.append((t, _lib.QjState.from_buffer_copy(s._st), s.t))
        return snap

    def _restore(self, snap):
        for s, (t, st, tt) in zip(self.sims, snap):
            for k, v in t.items():
                if k == "gas":
                    for a, b in zip(s.gas, v):
                        a.copy_(b)
                elif k == "gas_tmp":
                    for a, b in zip(s._gas_tmp, v):
                        a.copy_(b)
                elif v is not None:
                    getattr(s, k).copy_(v)
            C.memmove(C.addressof(s._st), C.addressof(st), C.sizeof(_lib.QjState))
            s.t = tt

    def _time(self, reps: int) -> float:
        self.step(1)  # records (and captures) the steady step
        torch.cuda.synchronize(self.device)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        self.step(reps)
        b.record()
        torch.cuda.synchronize(self.device)
        return a.elapsed_time(b) / reps

    # run-time schedule options the autotuner searches: (result key, option,
    # applies to the group of this shape key, candidate values).  The wide
    # sweep form exists for complex64 256-point x and y pencils; the z sweeps'
    # walking workgroups are a tile schedule of every group.
    KNOBS = (("fuse_x_source", _lib.OPT_FUSE_X_SOURCE, lambda key: True, (0, 1)),
             ("binned_deposit", _lib.OPT_BINNED_DEPOSIT, lambda key: key[2] == "stars", (0, 1)),
             ("star_gather", _lib.OPT_STAR_GATHER, lambda key: key[2] == "stars", (1, 0)),
             ("gather_rows", _lib.OPT_GATHER_ROWS, lambda key: key[2] == "stars", (0, 1, 2)),
             ("fused_deposit", _lib.OPT_FUSED_DEPOSIT, lambda key: key[2] == "stars", (2, 0)),
             ("wide_sweeps", _lib.OPT_WIDE_SWEEPS,
              lambda key: key[1] == "complex64" and 256 in key[0][:2], (0, 1)),
             # 65535: one pencil group per workgroup (no walking)
             ("z_sweep_blocks", _lib.OPT_Z_SWEEP_BLOCKS, lambda key: True,
              (0, 256, 1024, 2048, 65535)))

    def autotune(self, reps: int = 3, modes=("batched", "graph", "graph_fork"),
                 cache_path: str | None = None, verbose: bool = False) -> dict:
        """Pick the execution mode and, per shape group, the schedule options.

        Coordinate search: the mode is chosen with every group on its current
        options, then each group's options are set in turn to each of their
        candidate values and kept when the ensemble step gets faster (the
        options of KNOBS: the x-sweep schedule; for star groups the
        tile-binned vs float-atomic deposit, the fused star gather, its
        column tiles' rows and the fused deposit; for complex64 256-point
        groups the wide or narrow Poisson sweep form; the number of walking
        workgroups of the z sweeps).  Results are cached per ensemble
        signature in this process (and in cache_path, a JSON file, when
        given).  Returns {"mode", one {group: value} dict per knob name
        ("fuse_x_source", "binned_deposit", ...), "ms_per_step", "trials"}.
        """
        sig = self.signature()
        cache = dict(_TUNED)
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                cache.update(json.load(fh))
        groups = self.shape_groups()
        if sig in cache:
            self._apply(cache[sig], groups)
            return cache[sig]
        snap = self._snapshot()
        trials = []
        try:
            best_mode, best = None, float("inf")
            for m in modes:
                self.set_mode(m)
                ms = self._time(reps)
                trials.append({"mode": m, "ms": ms})
                if verbose:
                    print(f"[autotune] mode={m}: {ms:.3f} ms/step")
                if ms < best:
                    best_mode, best = m, ms
            self.set_mode(best_mode)
            chosen = {name: {} for name, _, _, _ in self.KNOBS}
            for name, opt, applies, values in self.KNOBS:
                for key, members in groups.items():
                    if not applies(key):
                        continue
                    plans = [self.sims[i].plan for i in members]
                    cur = plans[0].get_option(opt)
                    for val in values:
                        if val == cur:
                            continue
                        for p in plans:
                            p.set_option(opt, val)
                        ms = self._time(reps)
                        trials.append({"group": str(key), name: val, "ms": ms})
                        if verbose:
                            print(f"[autotune] {key} {name}={val}: {ms:.3f} ms/step")
                        if ms < best:
                            best, cur = ms, val
                    for p in plans:
                        p.set_option(opt, cur)
                    chosen[name][str(key)] = cur
        finally:
            self._restore(snap)
            torch.cuda.synchronize(self.device)
        res = {"mode": best_mode, **chosen, "ms_per_step": best, "trials": trials}
        _TUNED[sig] = res
        if cache_path:
            cache[sig] = res
            with open(cache_path, "w") as fh:
                json.dump(cache, fh, indent=1)
        return res

    def _apply(self, res: dict, groups: dict):
        self.set_mode(res["mode"])
        for name, opt, _, _ in self.KNOBS:
            for key, members in groups.items():
                v = res.get(name, {}).get(str(key))
                if v is not None:
                    for i in members:
                        self.sims[i].plan.set_option(opt, int(v))
