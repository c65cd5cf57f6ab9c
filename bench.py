"""Benchmark: cell-steps/s of the per-timestep hot path on one MI355X.

Workload: the 3x3 borehole array (401 x 401 x 1002 cells, FP64, dt 300 s,
per-well inlets) -- the plain-FP64 GPU configuration of the metric.  A step is
one full dt: ROCK2 z-step, explicit-y, advection, implicit-x + explicit-x,
advection, implicit-y.  The state stays resident in HBM.  The library runs
the steps as a fused pipeline (the implicit y of one step shares a kernel
with the next step's z-step and explicit y), as one captured hipGraph
replayed per step or as direct launches of the same kernels, whichever the
context's first long call measured faster (`config.steps_launch`).

Besides the headline line the JSON carries, as extra objects: `batched` (64
ragged single-borehole scenarios one at a time in FP64 against all of them in
one context in FP64, FP32 and FP32 with stochastic rounding), `large_fp32`
(the 5x5 array, 601 x 601 x 2002 cells, both FP32 states against FP64) and
`adjoint` (the layered single borehole,
121 x 121 x 602 cells: the 30-day outlet gradient's cost in forward steps and
central-difference checks); each can be switched off by its `--*-steps 0`.
`roofline` also gives the rate of a plain device copy of the field measured
on the same box (`copy_GBps`) and the dominant kernel's fraction of it.

Prints one JSON line (see the keys below).  `--gpus N` with N > 1 exits: this
build measures one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import __graft_entry__ as entry  # noqa: E402

HBM_PEAK = 8.0e12  # B/s, MI355X
METRIC = "cell-steps/s per GPU, plain FP64 vs batched/FP32 at ≤1e-3 K outlet err; %HBM BW"
# kernels of the fused step pipeline (gw_profile_steps): per step
# advect_1, implicit_explicit_x, advect_2 and implicit y fused with the next
# step's z-step and explicit y ("yzy"); the first step's z-step + explicit y
# ("zy") and the last step's implicit y ("implicit_y") run once.
OPS = ["zy", "advect_1", "implicit_explicit_x", "advect_2", "yzy", "implicit_y"]


def pipe_bytes(cache, w, s):
    """Algorithmic pipe bytes per step: 2 sub-steps x (advection read+write of
    every pipe cell, outlet read, turnaround annulus level means)."""
    bhs = w["boreholes"]
    inner, outer = cache.index_lists(0)
    n_in, n_an = len(inner), len(outer)
    gz = np.asarray(w["gridz"])
    c = w["dt"] / 2
    n_out = sum(int(np.sum(gz <= float(b.v_inner) * c)) for b in bhs) / len(bhs)
    countz = cache.countz
    return 2 * ((n_in + n_an) * countz * 2 * s + n_in * n_out * s + n_an * countz * s)


PROFILE = os.path.join(ROOT, "profiles", "c1_3x3_fp64_step_kernels.md")


def profile_traffic(mode):
    """HBM bytes of one application of the line operator `mode` from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload: its
    steady-state kernel (the k_line instance with the most calls) summed over
    the launch ranges it is split into (its rows with that call count, one per
    launch size; rows with fewer calls are the autotuner's timing launches of
    other variants), with the kernel name and the summed trace duration.  None
    if the summary is absent or not in the expected form."""
    if not os.path.exists(PROFILE):
        return None
    try:
        col = None
        rows = {}
        for line in open(PROFILE):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cells and cells[0] == "kernel":
                col = {name: q for q, name in enumerate(cells)}
                continue
            if col is None or len(cells) != len(col) or not cells[0].startswith("`void k_line<"):
                continue
            if not (cells[0].endswith(f"double, {mode}>`") or cells[0].endswith(f"double, {mode}, false>`")):
                continue
            name = cells[0].strip("`")
            b = (float(cells[col["HBM read MiB (scaled)"]]) + float(cells[col["HBM write MiB (scaled)"]])) * 2 ** 20
            rows.setdefault(name, []).append((int(cells[col["calls"]]), float(cells[col["avg us"]]), b))
        if not rows:
            return None
        name = max(rows, key=lambda n: max(r[0] for r in rows[n]))
        top = max(r[0] for r in rows[name])
        ranges = [r for r in rows[name] if r[0] == top]
        nbytes = sum(r[2] for r in ranges)
        if not nbytes > 0:
            return None
        return {"kernel": name, "trace_avg_us": sum(r[1] for r in ranges), "bytes_per_launch": nbytes}
    except (ValueError, KeyError, OSError):
        return None


def copy_rate(u, reps=5):
    """Bytes per second (read + write) of a plain device copy of the field
    on this box: the practical ceiling of a kernel that reads and writes the
    field once."""
    import torch
    tmp = torch.empty_like(u)
    tmp.copy_(u)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        tmp.copy_(u)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del tmp
    return 2 * u.numel() * u.element_size() / best


def _oracle_rate(w, threads, steps):
    """cell-steps/s of the oracle on `steps` steps of workload w from its
    initial field, with `threads` OpenMP threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers as H
    O = entry.load_oracle()
    O.set_threads(threads)
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    _, _, sec = orc.run(u0, 0.0, w["dt"], steps, 1.0 / 16.0)
    cells = len(w["gridx"]) * len(w["gridy"]) * len(w["gridz"])
    return cells * steps / sec, sec


def cpu_baseline(threads, c1_steps, c4_steps, single_thread):
    """The oracle -- a restatement of the reference CPU path that keeps its
    per-cell structure (point-wise material lookups over all boreholes,
    line-per-thread Thomas, the quadratic turnaround loop) -- timed on the
    host cores: the single borehole for its whole hour with `threads` threads
    and with one, one step of the 3x3 array (the headline workload: `value`)
    and one step of the 5x5 array.  Cost per step is constant over a run, so
    the rates are cell-steps/s."""
    from geothermalwells_jl_amd import workloads as W
    runs = []

    def run(name, w, thr, steps):
        rate, sec = _oracle_rate(w, thr, steps)
        runs.append({"workload": name, "threads": thr, "steps": steps, "seconds": sec, "cell_steps_per_s": rate})
        return rate

    c0 = W.c0()
    run("single borehole 41x41x202, whole hour", c0, threads, c0["nsteps"])
    if single_thread:
        run("single borehole 41x41x202, whole hour", c0, 1, c0["nsteps"])
    value = run("3x3 array 401x401x1002", W.c1(), threads, c1_steps)
    if c4_steps > 0:
        run("5x5 array 601x601x2002", W.c4(), threads, c4_steps)
    return dict(value=value, unit="cell-steps/s", cores=threads, kind="port",
                host_cpus=os.cpu_count(),
                sample=f"{c1_steps} step(s) of the 401x401x1002 workload with {threads} threads (the host's CPU "
                       f"share; os.cpu_count() = {os.cpu_count()}); OpenMP restatement of the reference CPU path "
                       f"in oracle/; the other runs are listed in `runs`",
                runs=runs)


def _init_state(gw, cache, ws, dtype):
    import torch
    u = cache.new_state()
    for s, w in enumerate(ws):
        gz = np.asarray(w["gridz"], dtype=np.float64)
        col = torch.from_numpy(w["ic"]["T_surface"] + w["ic"]["gradient"] * gz)
        cache.scenario_view(u, s).copy_(col.to(dtype=u.dtype).expand(*cache.scenario_shape(s)))
    return u


TUNE_STEPS = 9  # the shortest gw_steps call that times graph replays against direct launches


def _tune(gw, cache, u, dt):
    """The context's one-time choices (line-tile autotune, graph replays or
    direct launches) made by an untimed call on a scratch copy of the state,
    so the state itself gets exactly the warmup and timed steps asked for and
    no timing falls inside the timed region whatever the warmup count."""
    import torch
    tmp = u.clone()
    gw._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(tmp), 0.0, dt, TUNE_STEPS))
    cache.sync()
    torch.cuda.synchronize()
    del tmp


def _timed_steps(gw, cache, u, dt, warmup, steps):
    import torch
    lib = cache._lib
    _tune(gw, cache, u, dt)
    ptr = cache.check_state(u)
    gw._abi.check(lib.gw_steps(cache.ctx, ptr, 0.0, dt, warmup))
    cache.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gw._abi.check(lib.gw_steps(cache.ctx, ptr, warmup * dt, dt, steps))
    cache.sync()
    return time.perf_counter() - t0


def bench_batched(gw, steps, warmup, err_steps):
    """The 64 ragged single-borehole scenarios: one scenario at a time in FP64
    (a context each) vs all 64 in one context (every launch covers all of
    them) in FP64 and in FP32 state (rounded to nearest, and stochastically);
    the FP32 outlet errors are taken against the batched FP64 run of
    identical inputs after the same steps."""
    import torch
    from geothermalwells_jl_amd import workloads as W
    ws = W.c2()
    dt = ws[0]["dt"]
    cells = [W.cells(w) for w in ws]
    total = sum(cells)
    t_plain = 0.0
    for w in ws:
        cache = gw.create_cache(precision="f64", **W.scenario_kwargs(w))
        u = _init_state(gw, cache, [w], torch.float64)
        t_plain += _timed_steps(gw, cache, u, dt, warmup, steps)
        del cache, u
    res = {"workload": f"{len(ws)} ragged single-borehole scenarios "
                       f"({min(len(w['gridx']) for w in ws)}-{max(len(w['gridx']) for w in ws)} points in x and y, "
                       f"{min(len(w['gridz']) for w in ws)}-{max(len(w['gridz']) for w in ws)} in z), dt 120 s",
           "cells": total, "steps": steps}
    res["plain_fp64_one_at_a_time"] = total * steps / t_plain
    outs = {}
    n_err = max(err_steps, warmup + steps)
    for prec, dtype in (("f64", torch.float64), ("f32", torch.float32), ("f32sr", torch.float32)):
        cache = gw.create_batch_cache([W.scenario_kwargs(w) for w in ws], precision=prec)
        u = _init_state(gw, cache, ws, dtype)
        t = _timed_steps(gw, cache, u, dt, warmup, steps)
        res[f"batched_{prec}"] = total * steps / t
        rest = n_err - warmup - steps
        if rest > 0:
            gw._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), (warmup + steps) * dt, dt, rest))
        outs[prec] = cache.T_outlet
        del cache, u
    res["outlet_err_after_steps"] = n_err
    res["speedup_batched_fp32_vs_plain_fp64"] = res["batched_f32"] / res["plain_fp64_one_at_a_time"]
    res["speedup_batched_fp64_vs_plain_fp64"] = res["batched_f64"] / res["plain_fp64_one_at_a_time"]
    res["speedup_batched_fp32sr_vs_plain_fp64"] = res["batched_f32sr"] / res["plain_fp64_one_at_a_time"]
    res["outlet_max_abs_err_fp32_vs_fp64_K"] = float(np.abs(outs["f32"] - outs["f64"]).max())
    res["outlet_max_abs_err_fp32sr_vs_fp64_K"] = float(np.abs(outs["f32sr"] - outs["f64"]).max())
    res["unit"] = "cell-steps/s"
    return res


def bench_large(gw, steps, warmup, err_steps):
    """The 5x5 array (601 x 601 x 2002 cells, dt 600 s): FP32 state (FP64
    Thomas elimination and outlet reductions; rounded to nearest, and with
    stochastic rounding) against FP64, cell-steps/s of each and the FP32
    outlet errors after err_steps steps from the same initial field."""
    import torch
    from geothermalwells_jl_amd import workloads as W
    w = W.c4()
    cells = W.cells(w)
    dt = w["dt"]
    res = {"workload": "5x5 borehole array, 601x601x2002 cells, dt 600 s, per-well inlets", "cells": cells,
           "steps": steps, "unit": "cell-steps/s"}
    outs = {}
    n_err = max(err_steps, warmup + steps)
    for prec, dtype in (("f64", torch.float64), ("f32", torch.float32), ("f32sr", torch.float32)):
        cache = gw.create_cache(precision=prec, **W.scenario_kwargs(w))
        u = _init_state(gw, cache, [w], dtype)
        t = _timed_steps(gw, cache, u, dt, warmup, steps)
        res[prec] = cells * steps / t
        rest = n_err - warmup - steps
        if rest > 0:
            gw._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), (warmup + steps) * dt, dt, rest))
        outs[prec] = cache.T_outlet
        del cache, u
        torch.cuda.empty_cache()
    res["speedup_f32_vs_f64"] = res["f32"] / res["f64"]
    res["speedup_f32sr_vs_f64"] = res["f32sr"] / res["f64"]
    res["outlet_err_after_steps"] = n_err
    res["outlet_max_abs_err_fp32_vs_fp64_K"] = float(np.abs(outs["f32"] - outs["f64"]).max())
    res["outlet_max_abs_err_fp32sr_vs_fp64_K"] = float(np.abs(outs["f32sr"] - outs["f64"]).max())
    return res


def _objective(gw, cache, u0, dt, n):
    """J = mean over steps of the mean outlet, by the forward steps alone (one
    gw_steps call and one outlet read per step)."""
    import torch
    u = u0.clone()
    torch.cuda.synchronize()
    ptr = cache.check_state(u)
    acc = 0.0
    for s in range(n):
        gw._abi.check(cache._lib.gw_steps(cache.ctx, ptr, s * dt, dt, 1))
        acc += float(cache.T_outlet.mean())
    return acc / n


def bench_adjoint(gw, steps, seg, n_fd):
    """The layered single borehole (121 x 121 x 602 cells, 20 rock layers,
    dt 300 s): gradient of the mean outlet over `steps` steps with respect to
    the 20 layer conductivities and the inlet temperature, checkpointed every
    `seg` steps; its cost against the forward steps alone, and central
    differences (h = 1e-6 relative) of the inlet and the n_fd layers of
    largest |dJ/dk|."""
    import torch
    from geothermalwells_jl_amd import workloads as W
    w = W.c3()
    cells = W.cells(w)
    dt = w["dt"]
    kw = W.scenario_kwargs(w)
    cache = gw.create_cache(precision="f64", **kw)
    u0 = _init_state(gw, cache, [w], torch.float64)
    lib = cache._lib
    u = u0.clone()
    torch.cuda.synchronize()
    gw._abi.check(lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, dt, 9))  # tuning, graphs or direct
    cache.sync()
    u.copy_(u0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    gw._abi.check(lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, dt, steps))
    cache.sync()
    t_fwd = time.perf_counter() - t
    gw.outlet_gradient(cache, u0, 0.0, dt, 2, checkpoint_every=seg)
    torch.cuda.synchronize()
    t = time.perf_counter()
    J, dk, dT = gw.outlet_gradient(cache, u0, 0.0, dt, steps, checkpoint_every=seg)
    t_grad = time.perf_counter() - t
    res = {"workload": "single borehole, 20 rock layers, 121x121x602 cells, FP64, dt 300 s, constant inlet",
           "cells": cells, "steps": steps, "checkpoint_every": seg, "J_K": J,
           "forward_ms_per_step": 1e3 * t_fwd / steps, "gradient_ms_per_step": 1e3 * t_grad / steps,
           "gradient_cost_in_forward_steps": t_grad / t_fwd,
           "dJ_dk": [float(x) for x in dk], "dJ_dTin": float(dT.sum())}
    del cache, u
    mats = kw["materials"]
    fd = []
    h = 1e-6

    def J_of(materials=mats, T_in=None):
        k2 = dict(kw, materials=materials)
        if T_in is not None:
            k2["inlet_model"] = gw.ConstantInlet(T_in)
        c = gw.create_cache(precision="f64", **k2)
        return _objective(gw, c, u0, dt, steps)

    def layered(q, f):
        kl = list(mats.k_rock_layers)
        kl[q] *= f
        return gw.StratifiedMaterialProperties(kl, mats.rho_c_rock_layers, mats.layer_depths, mats.k_water,
                                               mats.rho_c_water, mats.k_steel, mats.rho_c_steel,
                                               mats.k_insulating, mats.rho_c_insulating, mats.k_backfill,
                                               mats.rho_c_backfill)

    T0 = kw["inlet_model"].T_const
    d = (J_of(T_in=T0 * (1 + h)) - J_of(T_in=T0 * (1 - h))) / (2 * h * T0)
    fd.append({"param": "T_in", "adjoint": float(dT.sum()), "central_difference": d})
    for q in np.argsort(-np.abs(dk))[:n_fd]:
        k0 = mats.k_rock_layers[q]
        d = (J_of(layered(int(q), 1 + h)) - J_of(layered(int(q), 1 - h))) / (2 * h * k0)
        fd.append({"param": f"k_layer_{int(q) + 1}", "adjoint": float(dk[q]), "central_difference": d})
    for e in fd:
        e["rel_diff"] = abs(e["adjoint"] - e["central_difference"]) / max(abs(e["central_difference"]), 1e-300)
    res["fd_check"] = fd
    torch.cuda.empty_cache()
    return res


def run_horizon(gw, name, precisions, dt, every, out_path):
    """A workload's whole simulated span (the 3x3 array: 30 days; the 64
    ragged scenarios: 7 days, batched; the 5x5 array: 1 year) in each
    precision, the contexts stepped in alternating chunks of `every` steps
    (default one simulated day).  At every chunk end: the outlet
    temperatures and the field's range.  Prints one JSON line with the
    largest and final outlet difference of each FP32 precision against FP64
    (when both ran) and writes the series to out_path (.npz) when given."""
    import torch
    from geothermalwells_jl_amd import workloads as W
    if name == "c2":
        ws = W.c2()
    else:
        ws = [getattr(W, name)(dt=dt) if (name == "c4" and dt) else getattr(W, name)()]
    dt = dt or ws[0]["dt"]
    span = {"c1": 30 * 86400.0, "c2": 7 * 86400.0, "c4": 365 * 86400.0}[name]
    nsteps = int(round(span / dt))
    every = every or int(round(86400.0 / dt))
    runs = {}
    for prec in precisions:
        dtype = torch.float64 if prec == "f64" else torch.float32
        cache = (gw.create_batch_cache([W.scenario_kwargs(w) for w in ws], precision=prec) if len(ws) > 1
                 else gw.create_cache(precision=prec, **W.scenario_kwargs(ws[0])))
        runs[prec] = dict(cache=cache, u=_init_state(gw, cache, ws, dtype), T=[], lo=[], hi=[])
    t_start = time.perf_counter()
    done = 0
    while done < nsteps:
        n = min(every, nsteps - done)
        for prec, r in runs.items():
            c = r["cache"]
            gw._abi.check(c._lib.gw_steps(c.ctx, c.check_state(r["u"]), done * dt, dt, n))
        for prec, r in runs.items():
            r["cache"].sync()
            r["T"].append(r["cache"].T_outlet)
            lo, hi = torch.aminmax(r["u"])
            r["lo"].append(float(lo))
            r["hi"].append(float(hi))
        done += n
        if done % (10 * every) == 0 or done == nsteps:
            msg = " ".join(f"{p}: outlet {r['T'][-1].min():.6f}..{r['T'][-1].max():.6f} field "
                           f"{r['lo'][-1]:.4f}..{r['hi'][-1]:.4f}" for p, r in runs.items())
            print(f"horizon {name} step {done}/{nsteps} ({done * dt / 86400:.1f} d, "
                  f"{time.perf_counter() - t_start:.0f} s): {msg}", file=sys.stderr, flush=True)
    res = {"horizon": name, "dt_s": dt, "steps": nsteps, "days": nsteps * dt / 86400, "sample_every": every,
           "seconds": time.perf_counter() - t_start}
    for prec, r in runs.items():
        T = np.array(r["T"])
        res[prec] = {"outlet_final": [float(x) for x in T[-1]], "field_min": min(r["lo"]), "field_max": max(r["hi"]),
                     "finite": bool(np.all(np.isfinite(T)))}
    for prec in runs:
        if prec != "f64" and "f64" in runs:
            d = np.abs(np.array(runs[prec]["T"]) - np.array(runs["f64"]["T"]))
            res[f"outlet_err_{prec}_vs_f64_max_K"] = float(d.max())
            res[f"outlet_err_{prec}_vs_f64_final_K"] = float(d[-1].max())
    if out_path:
        arrays = {"t": dt * every * np.arange(1, len(next(iter(runs.values()))["T"]) + 1)}
        for prec, r in runs.items():
            arrays[f"T_{prec}"] = np.array(r["T"])
            arrays[f"min_{prec}"] = np.array(r["lo"])
            arrays[f"max_{prec}"] = np.array(r["hi"])
        np.savez(out_path, **arrays)
    return res


def compare_horizons(path_a, path_b):
    """Outlet difference between each FP32 series (f32, f32sr) of one horizon
    file and the FP64 series of another (runs made in separate calls)."""
    a, b = np.load(path_a, allow_pickle=False), np.load(path_b, allow_pickle=False)
    Tb = b["T_f64"]
    if not np.array_equal(a["t"], b["t"]):
        raise ValueError("the two series are not sampled at the same times")
    res = {"samples": int(len(a["t"])), "t_end_s": float(a["t"][-1])}
    for prec in ("f32", "f32sr"):
        if f"T_{prec}" not in a:
            continue
        d = np.abs(a[f"T_{prec}"] - Tb)
        res[prec] = {"outlet_err_max_K": float(d.max()), "outlet_err_final_K": float(d[-1].max()),
                     "outlet_err_max_at_s": float(a["t"][int(d.max(axis=1).argmax())])}
    return res


def _cpu_share():
    """Threads for the CPU baseline: OMP_NUM_THREADS when the host sets it (a
    GPU host's share of its cores), else every core."""
    n = os.cpu_count() or 1
    try:
        return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))
    except ValueError:
        return n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=1, help="steps of the 3x3 array on the CPU")
    ap.add_argument("--cpu-c4-steps", type=int, default=1, help="steps of the 5x5 array on the CPU (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=_cpu_share(),
                    help="OpenMP threads of the CPU baseline (default: the host's CPU share, OMP_NUM_THREADS "
                         "when set, else os.cpu_count())")
    ap.add_argument("--no-cpu-single-thread", action="store_true", help="skip the 1-thread CPU run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=30,
                    help="timed steps of the 64-scenario batched comparison (0: skip)")
    ap.add_argument("--batched-err-steps", type=int, default=720,
                    help="steps after which the FP32 outlet error of the batched comparison is taken")
    ap.add_argument("--large-steps", type=int, default=20,
                    help="timed steps of the 601x601x2002 FP32-vs-FP64 comparison (0: skip)")
    ap.add_argument("--large-err-steps", type=int, default=144,
                    help="steps after which the FP32 outlet error of that comparison is taken")
    ap.add_argument("--adjoint-steps", type=int, default=8640,
                    help="steps of the layered-borehole gradient (30 days at dt 300 s; 0: skip)")
    ap.add_argument("--adjoint-checkpoint", type=int, default=64, help="checkpoint interval of the gradient")
    ap.add_argument("--adjoint-fd", type=int, default=2,
                    help="layer conductivities checked by central differences (besides the inlet)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--horizon", choices=["c1", "c2", "c4"],
                    help="instead of the benchmark: run the workload's whole simulated span (3x3 array, 64 "
                         "scenarios batched, 5x5 array) and print its outlet errors")
    ap.add_argument("--horizon-precision", default="f64,f32", help="precisions of --horizon (comma separated)")
    ap.add_argument("--horizon-dt", type=float, default=0.0, help="dt of --horizon c4 (default: 600 s)")
    ap.add_argument("--horizon-every", type=int, default=0, help="outlet sample interval in steps (default: a day)")
    ap.add_argument("--horizon-out", default="", help="write the --horizon series to this .npz")
    ap.add_argument("--horizon-compare", nargs=2, metavar=("F32_NPZ", "F64_NPZ"),
                    help="compare two --horizon-out files (no GPU)")
    args = ap.parse_args()
    if args.horizon_compare:
        print(json.dumps(compare_horizons(*args.horizon_compare)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the step pipeline runs on a single device (a field split over "
              "GPUs would shard as z-slabs with a 3-level halo exchange per step, not implemented); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    gw = entry.load_package()
    if args.horizon:
        res = run_horizon(gw, args.horizon, args.horizon_precision.split(","), args.horizon_dt,
                          args.horizon_every, args.horizon_out)
        print(json.dumps(res), flush=True)
        return
    from geothermalwells_jl_amd import workloads as W
    w = W.c1()
    cells = W.cells(w)
    cache = gw.create_cache(precision="f64", **W.scenario_kwargs(w))
    lib = cache._lib
    u = gw.initial_condition_thermal_gradient("cuda", np.float64, w["gridx"], w["gridy"], w["gridz"],
                                              **w["ic"]).reshape(-1).contiguous()
    ptr = cache.check_state(u)
    dt = w["dt"]
    torch.cuda.synchronize()
    _tune(gw, cache, u, dt)
    gw._abi.check(lib.gw_steps(cache.ctx, ptr, 0.0, dt, args.warmup))
    cache.sync()
    torch.cuda.synchronize()

    stream = torch.cuda.ExternalStream(lib.gw_stream(cache.ctx))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start = time.perf_counter()
    e0.record(stream)
    gw._abi.check(lib.gw_steps(cache.ctx, ptr, args.warmup * dt, dt, args.steps))
    e1.record(stream)
    cache.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ev_ms = e0.elapsed_time(e1)
    ms_per_step = 1e3 * wall / args.steps
    value = cells * args.steps / wall
    T_out = cache.T_outlet

    # per-operator HIP-event timing on the context's stream, after a few
    # steps queued right before it: the GPU lowers its clocks within
    # milliseconds of idling (the host's checks above), and the first
    # launches after an idle gap would run the fused y kernel 10-20% slower
    ms = (ctypes.c_double * 6)()
    t0 = (args.warmup + args.steps) * dt
    gw._abi.check(lib.gw_steps(cache.ctx, ptr, t0, dt, 6))
    gw._abi.check(lib.gw_profile_steps(cache.ctx, ptr, t0 + 6 * dt, dt, args.profile_steps, ms))
    assert np.all(np.isfinite(T_out)), "non-finite outlet temperatures"
    # finiteness of the state through its sum (NaN and Inf propagate; an
    # elementwise isfinite() would allocate ~1.5 GB of temporaries)
    state_ok = bool(np.isfinite(float(u.sum().item())))
    n_p = args.profile_steps
    launches = [1, n_p, n_p, n_p, n_p - 1, 1]
    per_op = {OPS[q]: ms[q] / launches[q] for q in range(6) if launches[q] > 0}
    s = 8
    field_ops = ["implicit_explicit_x", "yzy"]
    dom = max(field_ops, key=lambda k: per_op[k])
    dom_bytes = cells * 2 * s  # one read + one write of the state per launch
    achieved = dom_bytes / (per_op[dom] * 1e-3)
    traffic = profile_traffic({"implicit_explicit_x": 2, "yzy": 4}[dom])
    copy_bw = copy_rate(u)
    B_cs = 10 * s
    Bp = pipe_bytes(cache, w, s)
    steps_per_s = args.steps / wall
    step_bw = value * B_cs + steps_per_s * Bp
    fused_bw = value * 4 * s + steps_per_s * Bp

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded inputs of the 3x3-array configuration; linear geothermal initial field)",
        "config": {"workload": "3x3 borehole array, 401x401x1002 cells, FP64, dt 300 s, per-well inlets",
                   "cells": cells, "boreholes": len(w["boreholes"]), "dt_s": dt,
                   "steps_launch": cache.steps_launch, "line_launch": cache.line_launch},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": (f"{os.path.relpath(PROFILE, ROOT)}: {traffic['kernel']} "
                                        f"(trace avg {traffic['trace_avg_us']:.1f} us)") if traffic else None,
                     "algorithmic_bytes_per_launch": dom_bytes, "avg_launch_ms": per_op[dom],
                     "copy_GBps": copy_bw / 1e9, "frac_of_copy": achieved / copy_bw},
        "step_roofline": {"B_cs_5sweep": B_cs, "pipe_bytes_per_step": Bp,
                          "frac_5sweep": step_bw / HBM_PEAK, "frac_2sweep_as_run": fused_bw / HBM_PEAK},
        "per_op_ms": per_op,
        "event_ms_per_step": ev_ms / args.steps,
        "outlet_T": [float(x) for x in T_out],
        "state_finite": state_ok,
    }
    if args.batched_steps > 0:
        del cache, u
        torch.cuda.empty_cache()
        out["batched"] = bench_batched(gw, args.batched_steps, 3, args.batched_err_steps)
    if args.large_steps > 0:
        if "cache" in locals():
            del cache, u
        torch.cuda.empty_cache()
        out["large_fp32"] = bench_large(gw, args.large_steps, 2, args.large_err_steps)
    if args.adjoint_steps > 0:
        if "cache" in locals():
            del cache, u
        torch.cuda.empty_cache()
        out["adjoint"] = bench_adjoint(gw, args.adjoint_steps, args.adjoint_checkpoint, args.adjoint_fd)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_threads, args.cpu_steps, args.cpu_c4_steps,
                                           not args.no_cpu_single_thread)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
