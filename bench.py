#!/usr/bin/env python3
"""Throughput of the HERCULES time step on one MI355X.

Workload (default ``fd``): the finite-difference RK3 step of a neutral
closed channel (cds=2: 4th-order horizontal, 2nd-order vertical; dp_opt=1;
Re_tau=180; 4pi x 2pi x 2; zstretch=2.0) at 256 x 256 x 257 cells, FP64,
on a seeded synthetic state (Reichardt log-law mean + 10% Gaussian noise).
``fd_t`` adds the temperature equation (stratified open channel, got=0.1);
``poisson`` times the standalone pressure Poisson solve (seeded Gaussian
right-hand side, plane means removed).

``--workload adjoint`` times the gradient of the wall shear after --steps
steps by the discrete adjoint (checkpoint every --ckpt steps) against the
forward run, on 128x128x129.

Before timing, hc_autotune picks the launch shapes (bit-identical results,
except the x-y tiled momentum shapes, which agree with the fused kernel to
round-off; ``--no-autotune`` keeps the defaults, ``--launch-shapes`` sets
given ones).

``--precision mixed`` times the mixed variant (FP32 fields, stencils and
FFTs; FP64 low-wavenumber column solves and bulk-flow update) instead.  For
fd, fd_t, sp and poisson the other precision is also run on the same state for
the same number of steps: the line then carries the FP64/mixed speed ratio
and the mixed variant's relative L2 error, wall-shear error and divergence
against FP64 (``mixed`` object), unless ``--no-compare``.

Prints one JSON line.  ``value`` is whole-job Mcell-RK3-steps/s
(nx*ny*nz*steps / seconds / 1e6) with the state resident in HBM; the
``roofline`` object is computed for the kernel with the largest share of
the step time from its algorithmic bytes per launch (fields it must read
and write, FP64) and its average duration measured with HIP events on the
library's stream during the timed region; ``cpu_baseline`` times the CPU
oracle (oracle/hc_oracle.c, OpenMP) on the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "Mcell-RK3-steps/s @256²×257: FP64 vs FP32+FP64-refine; adjoint/forward ratio"


def algo_passes(kernel, substep, with_t):
    """Algorithmic FP64 field passes of one launch (reads + writes of whole
    fields, a half spectrum counting as one field)."""
    if kernel == "rhs":
        reads = 4 + (1 if with_t else 0) + (3 if substep > 1 else 0)
        return reads + 6
    return {"yfft_fwd": 6, "column": 7, "yfft_inv": 8, "finish": 9,
            "rhs_t": None, "yfft_t": 2, "column_t": 2, "finish_t": 2,
            "solve_xr2c": 2, "solve_yfft": 2, "solve_column": 2, "solve_xc2r": 2}[kernel]


def rhs_t_passes(substep):
    return 4 + (1 if substep > 1 else 0) + 2


# Spectral substep (cds=3, skew-symmetric form, temperature), field passes per
# launch slot: rhs = y inverse of the six derivative spectra (read c_u, c_v,
# c_w, write 6) + products (read u, v, w, t and the 6, write 8); yfft_fwd = 8
# spectra in place; column = read 8 products, c_u, c_v, c_w, c_p and the old
# c_ff, c_gg, c_hh, write the 7 updated; finish = y and x inverses of c_u,
# c_v, c_w; rhs_t = read u, v, w, t, write 3; yfft_t = 3 in place; column_t =
# read 3, c_t, c_ss, write c_t, c_ss; finish_t = inverse of c_t.
SPECTRAL_PASSES = {"rhs": 27, "yfft_fwd": 16, "column": 22, "finish": 12, "rhs_t": 7, "yfft_t": 6,
                   "column_t": 7, "finish_t": 4}


def build_case(workload):
    from hercules_amd import cases
    # wall_rows=1: every mode gets the implicit wall rows that the reference
    # gives the (0,0) mode only; with the reference's rows this 257-level
    # grid is unstable (rk_c*dt*nu/(2 dz_min^2) > 1).  Same kernels, same work.
    if workload == "fd":
        prm = cases.closed_channel(nx=256, ny=256, nz=257, wall_rows=1)
        st = cases.loglaw_noise(prm, seed=2024)
        st.pop("t")
    elif workload == "fd_t":
        prm = cases.stratified_open_channel(nx=256, ny=256, nz=257, wall_rows=1)
        st = cases.loglaw_noise(prm, seed=2024)
    elif workload == "sp":
        # stratified spectral case with a constant temperature gradient at the top
        prm = cases.stratified_open_channel(nx=256, ny=256, nz=257, cds=3, issk=1, wall_rows=1, tbc_top=1,
                                            dtdz_top=1.0)
        st = cases.loglaw_noise(prm, seed=2024)
    elif workload == "poisson":
        prm = cases.closed_channel(nx=256, ny=256, nz=257)
        st = {"rhs": cases.gaussian_poisson_rhs(prm, seed=2024)}
    else:
        raise ValueError(workload)
    return prm, st


def ensemble_members(seed=2024):
    """The 24-member ensemble of hercules_amd.cases.ensemble_24."""
    from hercules_amd import cases
    return cases.ensemble_24(seed)


def run_ensemble(args):
    """Aggregate throughput of the 24-channel ensemble (one captured
    graph per two steps, members concurrent) against the same members
    stepped one after another."""
    from hercules_amd import Channel, Ensemble
    members = []
    cells = 0
    mixed = args.precision == "mixed"
    nmixed = 0
    for prm, st in ensemble_members():
        # the mixed variant has no Rayleigh damping: sponge members stay FP64
        pr = 1 if (mixed and prm.isdamp == 0) else 0
        nmixed += pr
        ch = Channel(prm.replace(precision=pr))
        for k, v in st.items():
            ch.set_field(k, v)
        ch.project()
        ch.init()
        if args.autotune:
            ch.autotune(steps_per_trial=1)
        members.append(ch)
        cells += prm.nx * prm.ny * prm.nz
    # sequential single-instance launches
    for ch in members:
        ch.steps(args.warmup)
        ch.synchronize()
    t0 = time.perf_counter()
    for ch in members:
        ch.steps(args.steps)
        ch.synchronize()
    seq = time.perf_counter() - t0
    ens = Ensemble(members, use_graph=True)
    ens.steps(2 * ((args.warmup + 1) // 2))
    t0 = time.perf_counter()
    ens.steps(args.steps)
    par = time.perf_counter() - t0
    graph = ens.graph_used
    finite = all(bool(np.isfinite(ch.get_field("u")).all()) for ch in members)
    ens.close()
    for ch in members:
        ch.close()
    value = cells * args.steps / par / 1e6
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(par / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (seeded log-law + noise states; no checkpoint)",
        "config": {"workload": "ensemble of 24 channels (8 x 128x128x129, 8 x 192x192x129, 8 x 256x128x193; neutral, stratified, Ekman-type), one HIP graph per two steps",
                   "members": len(members), "cells": cells, "precision": args.precision,
                   "mixed_members": nmixed},
        "sequential_value": round(cells * args.steps / seq / 1e6, 2),
        "speedup_vs_sequential": round(seq / par, 3), "graph_used": graph, "state_finite": finite,
    }
    print(json.dumps(out))


def run_adjoint(args):
    """Gradient of the bottom wall shear after K = --steps
    RK3 steps (128x128x129 neutral closed channel, Re_tau = 180) with respect
    to the initial velocity, by the hand-written adjoint with a checkpoint
    every --ckpt steps; against the forward run of the same K steps."""
    r = adjoint_case(args.steps, args.ckpt)
    K, n, t_grad = r.pop("K"), r.pop("cells"), r.pop("t_grad")
    value = n * K / t_grad / 1e6
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-RK3-steps/s (gradient)", "n_gpus": 1,
        "steps": K, "warmup": 1, "ms_per_step": round(t_grad / K * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded log-law + 5% noise, projected; no checkpoint file)",
        "config": r.pop("config"),
    }
    out.update(r)
    print(json.dumps(out))


def adjoint_case(K, ck):
    """The adjoint workload (128x128x129, gradient of the wall shear after K
    steps, checkpoint every ck): gradient and forward times, one central
    difference check, the HBM held."""
    from hercules_amd import Channel, cases
    prm = cases.closed_channel(nx=128, ny=128, nz=129, wall_rows=1)
    st = cases.loglaw_noise(prm, seed=2024, rel_amp=0.05)
    st.pop("t")
    n = prm.nx * prm.ny * prm.nz
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    x0 = {k: ch.get_field(k) for k in ("u", "v", "w", "p")}
    ch.init()
    ch.wall_shear_gradient(min(K, ck), ck)  # warm-up
    ch.synchronize()
    # forward K steps
    t0 = time.perf_counter()
    ch.steps(K)
    ch.synchronize()
    t_fwd = time.perf_counter() - t0
    j_fwd = ch.wall_shear()[0]
    for k, v in x0.items():
        ch.set_field(k, v)
    ch.init()
    t0 = time.perf_counter()
    loss = ch.wall_shear_gradient(K, ck)
    ch.synchronize()
    t_grad = time.perf_counter() - t0
    g = {k: ch.get_adjoint(k) for k in ("u", "v", "w")}
    # one central-difference check along a seeded divergence-free direction
    rng = np.random.default_rng(7)
    for k in ("u", "v", "w"):
        a = rng.standard_normal(x0[k].shape)
        if k == "w":
            a[-1] = 0.0
        ch.set_field(k, a)
    ch.project()
    d = {k: ch.get_field(k) for k in ("u", "v", "w")}
    eps = 1e-6 * np.sqrt(sum(np.vdot(x0[k], x0[k]) for k in d) / sum(np.vdot(d[k], d[k]) for k in d))
    js = []
    for sgn in (1.0, -1.0):
        for k, v in x0.items():
            ch.set_field(k, v + sgn * eps * d[k] if k in d else v)
        ch.init()
        ch.steps(K)
        js.append(ch.wall_shear()[0])
    fd = (js[0] - js[1]) / (2 * eps)
    ad = sum(float(np.vdot(g[k], d[k])) for k in d)
    ch.close()
    nseg = (K + ck - 1) // ck
    ckpt_bytes = (nseg * 4 + 3 * min(ck, K) * 4) * n * 8
    ctx_bytes = (13 + 13) * n * 8 + 4 * (prm.nx // 2) * prm.ny * prm.nz * 16
    return {
        "K": K, "cells": n, "t_grad": t_grad,
        "config": {"workload": f"wall-shear gradient after {K} RK3 steps, adjoint with a checkpoint every {ck} steps, "
                               "neutral closed channel Re_tau=180, 128x128x129, cds=2, wall_rows=1",
                   "nx": prm.nx, "ny": prm.ny, "nz": prm.nz, "cells": n, "precision": "fp64"},
        "gradient_ms_per_step": round(t_grad / K * 1e3, 4),
        "forward_ms_per_step": round(t_fwd / K * 1e3, 4),
        "adjoint_over_forward": round(t_grad / t_fwd, 3),
        "backward_over_forward": round((t_grad - 2 * t_fwd) / t_fwd, 3),
        "loss": loss, "loss_forward_run": j_fwd,
        "fd_check": {"eps_rel": 1e-6, "fd": fd, "adjoint": ad, "rel_err": abs(fd - ad) / abs(ad)},
        "peak_hbm_bytes_estimate": int(ckpt_bytes + ctx_bytes),
    }


def config3_case(steps, warmup, precision=0):
    """The spectral step of the stratified open channel with temperature,
    6th-order z operators and a constant top temperature gradient at
    256x256x257 (cds=3, skew-symmetric, 2/3 rule; Re_tau=395, got=0.1):
    time per step, kernel timings and the roofline of the dominant kernel."""
    from hercules_amd import Channel
    prm, st = build_case("sp")
    prm = prm.replace(zorder=6, precision=precision)
    n = prm.nx * prm.ny * prm.nz
    ch = Channel(prm.replace(precision=0))
    for k, v in st.items():
        ch.set_field(k, v)
    ch.project()
    for k in ("u", "v", "w"):
        st[k] = ch.get_field(k)
    ch.close()
    ch = Channel(prm)
    for k, v in st.items():
        ch.set_field(k, v)
    ch.init()
    ch.steps(warmup)
    ch.synchronize()
    ch.profile(True)
    t0 = time.perf_counter()
    ch.steps(steps)
    ch.synchronize()
    el = time.perf_counter() - t0
    prof = ch.profile_read()
    ch.profile(False)
    fields = {k: ch.get_field(k) for k in ("u", "v", "w", "t")}
    div = ch.screen_div_error()
    tau = ch.wall_shear()[0]
    ch.close()
    eb = 4 if precision else 8
    kernels = {}
    for name, (ms, launches) in prof.items():
        b = SPECTRAL_PASSES[name] * n * eb
        kernels[name] = {"avg_ms": round(ms / launches, 4), "launches": launches,
                         "GBs": round(b / (ms / launches * 1e-3) / 1e9, 1), "algo_bytes": b}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    kd = kernels[dom]
    ms_step = el / steps * 1e3
    step_gbs = 3 * sum(SPECTRAL_PASSES.values()) * n * eb / (ms_step * 1e-3) / 1e9
    return {"workload": "spectral RK3 step (cds=3, skew-symmetric, 2/3 rule) with temperature, 6th-order z "
                        "operators, top dT/dz=1, stratified open channel, Re_tau=395, got=0.1, 256x256x257",
            "precision": "mixed" if precision else "fp64", "steps": steps, "warmup": warmup,
            "ms_per_step": round(ms_step, 4), "value": round(n * steps / el / 1e6, 2), "unit": "Mcell-RK3-steps/s",
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": kd["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(kd["GBs"] / HBM_PEAK_GBS, 4), "avg_ms": kd["avg_ms"],
                         "algo_bytes_per_launch": int(kd["algo_bytes"])},
            "step_roofline": {"passes_per_step": 3 * sum(SPECTRAL_PASSES.values()), "achieved": round(step_gbs, 1),
                              "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
            "kernels": {k: {kk: vv for kk, vv in v.items() if kk != "algo_bytes"} for k, v in kernels.items()},
            "state_finite": bool(all(np.isfinite(f).all() for f in fields.values())), "div": div,
            "wall_shear_bottom": tau, "_fields": fields}


def cpu_baseline(prm, st, workload, steps):
    """Oracle on this host's cores: `steps` full steps (or solves) of the same
    workload, start-up excluded."""
    from oracle.oracle import Oracle, lib
    from tests.helpers import oracle_kwargs
    o = Oracle(**oracle_kwargs(prm))
    threads = lib().ho_num_threads()
    n = prm.nx * prm.ny * prm.nz
    if workload == "poisson":
        o.init()
        t0 = time.perf_counter()
        for _ in range(steps):
            o.solve("p", 1.0, st["rhs"])
        dt = time.perf_counter() - t0
        val, unit = steps / dt, "Poisson solves/s"
        sample = f"{steps} pressure solve(s) at {prm.nx}x{prm.ny}x{prm.nz}"
    else:
        for k, v in st.items():
            o.set(k, v)
        o.init()
        t0 = time.perf_counter()
        o.steps(steps)
        dt = time.perf_counter() - t0
        val, unit = n * steps / dt / 1e6, "Mcell-RK3-steps/s"
        sample = f"{steps} RK3 step(s) of the same {prm.nx}x{prm.ny}x{prm.nz} workload, start-up excluded"
    o.close()
    return {"value": val, "unit": unit, "cores": int(threads), "kind": "port", "sample": sample,
            "seconds": dt}


def copy_calibration(args):
    """Plain device-to-device copy of one 256x256x257 FP64 field (torch's copy
    kernel): the known-size reference for the FETCH_SIZE / WRITE_SIZE
    counters and the attainable HBM bandwidth of a streaming kernel."""
    import torch
    n = 256 * 256 * 257
    a = torch.randn(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2 * n * 8 / (ms * 1e-3) / 1e9
    print(json.dumps({"workload": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "ms": round(ms, 4), "GBs": round(gbs, 1)}))


class DeviceArray:
    """An FP64 array in device memory (hipMalloc through ctypes)."""

    _hip = None

    def __init__(self, host):
        import ctypes as C
        if DeviceArray._hip is None:
            for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
                try:
                    DeviceArray._hip = C.CDLL(name)
                    break
                except OSError:
                    continue
        H = DeviceArray._hip
        self.shape, self.nbytes = host.shape, host.nbytes
        p = C.c_void_p()
        if H.hipMalloc(C.byref(p), C.c_size_t(self.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr = p.value
        a = np.ascontiguousarray(host, dtype=np.float64)
        if H.hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), C.c_size_t(self.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy to device failed")

    def get(self):
        import ctypes as C
        out = np.empty(self.shape)
        if DeviceArray._hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr),
                                      C.c_size_t(self.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def __del__(self):
        import ctypes as C
        if getattr(self, "ptr", None) and DeviceArray._hip is not None:
            DeviceArray._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None


def timed_run(prm, st, args, precision, profile):
    """Warm up, then time args.steps steps (or solves) of one channel in the
    given precision (0 FP64, 1 mixed), state resident on the device."""
    from hercules_amd import Channel
    ch = Channel(prm.replace(precision=precision))
    if args.workload == "poisson":
        rhs = np.ascontiguousarray(st["rhs"])
        # device-resident FP64 source and result: two device arrays of the
        # HIP runtime; the solver reads and writes them through the C ABI
        src, dst = DeviceArray(rhs), DeviceArray(np.zeros_like(rhs))
        step = lambda: ch.poisson_solver_fft_device("p", 1.0, src.ptr, dst.ptr)
    else:
        for k, v in st.items():
            ch.set_field(k, v)
        ch.init()
        if getattr(args, "launch_shapes", None):
            ch.set_tuning(**args.launch_shapes)
            ch.tuning = dict(args.launch_shapes)
        elif getattr(args, "autotune", False):
            ch.tuning = ch.autotune(steps_per_trial=2)
        step = lambda: ch.steps(1)
    for _ in range(args.warmup):
        step()
    ch.synchronize()
    if profile:
        ch.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ch.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ch.profile_read() if profile else {}
    ch.profile(False)
    res = {"ch": ch, "elapsed": elapsed, "prof": prof, "finite": True, "div": None}
    if args.workload == "poisson":
        res["phi"] = dst.get()
        res["rhs"] = rhs
        res["bufs"] = (src, dst)
    else:
        names = ["u", "v", "w", "p"] + (["t"] if prm.isscalar else [])
        res["fields"] = {k: ch.get_field(k) for k in names}
        res["finite"] = bool(all(np.isfinite(f).all() for f in res["fields"].values()))
        try:
            res["div"] = ch.screen_div_error()
        except Exception as e:  # NaN in the state
            res["div"] = str(e)
        res["tau"] = ch.wall_shear()
    return res


def compare_precisions(prm, args, m, d, n):
    """Mixed run m against FP64 run d after the same steps from the same state."""
    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    ms_m = m["elapsed"] / args.steps * 1e3
    ms_d = d["elapsed"] / args.steps * 1e3
    out = {"ms_per_step_mixed": round(ms_m, 4), "ms_per_step_fp64": round(ms_d, 4),
           "mixed_speedup_vs_fp64": round(ms_d / ms_m, 3)}
    if args.workload == "poisson":
        out["value_mixed"] = round(args.steps / m["elapsed"], 2)
        out["unit"] = "Poisson solves/s"
        out["rel_l2_phi"] = rel(m["phi"], d["phi"])
    else:
        out["value_mixed"] = round(n * args.steps / m["elapsed"] / 1e6, 2)
        out["unit"] = "Mcell-RK3-steps/s"
        out["after_steps"] = args.warmup + args.steps
        out["rel_l2"] = {k: rel(m["fields"][k], d["fields"][k]) for k in d["fields"]}
        tb_m, tb_d = m["tau"][0], d["tau"][0]
        out["wall_shear_bottom"] = {"fp64": tb_d, "mixed": tb_m, "rel_err": abs(tb_m - tb_d) / abs(tb_d)}
        out["div_fp64"] = d["div"]
        out["div_mixed"] = m["div"]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="fd", choices=["fd", "fd_t", "sp", "poisson", "ensemble", "adjoint",
                                                          "copy"])
    ap.add_argument("--ckpt", type=int, default=20, help="adjoint workload: steps between checkpoints")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shapes (otherwise hc_autotune picks them before timing; "
                         "fd, fd_t, ensemble)")
    ap.add_argument("--zorder", type=int, default=2, choices=[2, 4, 6],
                    help="order of the z operators (4, 6: the high-order extension; fd, fd_t, sp, poisson)")
    ap.add_argument("--launch-shapes", default=None,
                    help="fixed launch shapes instead of autotuning, e.g. rgroup=16,ytile=8,c2r_pts=1024,col_remap=1,rhs_xt=2 "
                         "(the config.launch_shapes an autotuned run printed; fd, fd_t, sp)")
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle steps for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event timing")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="fp64, or FP32 fields/stencils/FFTs with FP64 low-wavenumber solves")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the run in the other precision (speed ratio and mixed-vs-FP64 errors)")
    ap.add_argument("--no-extras", action="store_true",
                    help="fd: skip the adjoint (128x128x129, 400 steps, checkpoint every 20) and the "
                         "6th-order-z spectral step (256x256x257) that the default line also carries")
    args = ap.parse_args()
    if args.zorder != 2 and args.workload in ("ensemble", "adjoint", "copy"):
        ap.error("--zorder 4/6 covers fd, fd_t, sp and poisson")
    if args.launch_shapes:
        try:
            args.launch_shapes = {k: int(v) for k, v in (kv.split("=") for kv in args.launch_shapes.split(","))}
        except ValueError:
            ap.error("--launch-shapes takes knob=value pairs separated by commas")
        unknown = set(args.launch_shapes) - {"rgroup", "ytile", "c2r_pts", "col_remap", "rhs_xt"}
        if unknown:
            ap.error(f"unknown launch-shape knobs {sorted(unknown)}")
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the channel is resident in one MI355X's HBM)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        return copy_calibration(args)
    if args.workload == "ensemble":
        return run_ensemble(args)
    if args.workload == "adjoint":
        return run_adjoint(args)

    prm, st = build_case(args.workload)
    if args.zorder != 2:
        prm = prm.replace(zorder=args.zorder)
    n = prm.nx * prm.ny * prm.nz
    with_t = args.workload == "fd_t"
    mixed_main = args.precision == "mixed"
    if args.workload != "poisson":
        from hercules_amd import Channel
        ch0 = Channel(prm)
        for k, v in st.items():
            ch0.set_field(k, v)
        ch0.project()  # seeded noise made discretely divergence-free (FP64)
        for k in ("u", "v", "w"):
            st[k] = ch0.get_field(k)
        ch0.close()
    run = timed_run(prm, st, args, 1 if mixed_main else 0, profile=not args.no_profile)
    ch, elapsed, prof = run["ch"], run["elapsed"], run["prof"]
    ms_per_step = elapsed / args.steps * 1e3
    finite, div = run["finite"], run["div"]
    # the other precision on the same state and step count: its speed and
    # the mixed variant's departure from FP64
    other = None
    if args.workload in ("fd", "fd_t", "sp", "poisson") and not args.no_compare:
        alt = timed_run(prm, st, args, 0 if mixed_main else 1, profile=False)
        m, d = (run, alt) if mixed_main else (alt, run)
        other = compare_precisions(prm, args, m, d, n)
        alt["ch"].close()

    if args.workload == "poisson":
        value = args.steps / elapsed
        unit = "Poisson solves/s"
    else:
        value = n * args.steps / elapsed / 1e6
        unit = "Mcell-RK3-steps/s"

    # per-kernel algorithmic bandwidth
    kernels = {}
    nsub = 3 if prm.dts == 2 else 1
    for name, (ms, launches) in prof.items():
        if args.workload == "sp":
            passes = SPECTRAL_PASSES[name]
        elif name == "rhs":
            per_step = sum(algo_passes("rhs", s, with_t) for s in range(1, nsub + 1))
            passes = per_step / nsub
        elif name == "rhs_t":
            passes = sum(rhs_t_passes(s) for s in range(1, nsub + 1)) / nsub
        else:
            passes = algo_passes(name, 1, with_t)
        bytes_per_launch = passes * n * (4 if mixed_main else 8)
        avg_ms = ms / launches
        kernels[name] = {"avg_ms": avg_ms, "launches": launches, "share": 0.0,
                         "algo_bytes": bytes_per_launch,
                         "GBs": bytes_per_launch / (avg_ms * 1e-3) / 1e9}
    tot = sum(v["avg_ms"] * v["launches"] for v in kernels.values()) or 1.0
    for v in kernels.values():
        v["share"] = v["avg_ms"] * v["launches"] / tot
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE runs,
    # calibrated against a copy kernel), committed under profiles/
    traffic = {}
    tag = args.workload + ("_mixed" if mixed_main else "") + (f"_z{args.zorder}" if args.zorder != 2 else "")
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}_{prm.nx}x{prm.ny}x{prm.nz}.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic = dict(tj.get("bytes_per_launch", {}))
        # the right-hand-side stage's bytes depend on the momentum tile shape
        shapes = getattr(ch, "tuning", None) or {}
        by_shape = tj.get("rhs_by_shape", {})
        if str(shapes.get("rhs_xt", 0)) in by_shape:
            traffic["rhs"] = by_shape[str(shapes.get("rhs_xt", 0))]
        elif "launch_shapes" in tj and shapes.get("rhs_xt", 0) != tj["launch_shapes"].get("rhs_xt", 0):
            traffic.pop("rhs", None)
    roofline = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
        kd = kernels[dom]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(kd["GBs"], 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(kd["GBs"] / HBM_PEAK_GBS, 4),
                    "traffic": traffic.get(dom), "avg_ms": round(kd["avg_ms"], 4),
                    "algo_bytes_per_launch": int(kd["algo_bytes"])}
    # whole step against the step's algorithmic bytes (1272 B/cell/step for FD
    # without T, 1608 with T: 159 / 201 FP64 passes per RK3 step)
    step_passes = {"fd": 159, "fd_t": 201, "sp": 3 * sum(SPECTRAL_PASSES.values()),
                   "poisson": 10}[args.workload]
    step_gbs = step_passes * n * (4 if mixed_main else 8) / (ms_per_step * 1e-3) / 1e9
    step_roofline = {"passes_per_step": step_passes, "achieved": round(step_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    cpu = None
    if not args.no_cpu_baseline and args.zorder != 2:
        cpu = {"value": None, "note": "the oracle restates the reference, which is 2nd order in z"}
    elif not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(prm, st, args.workload, args.cpu_steps)
        except Exception as e:  # the baseline never gates the GPU number
            cpu = {"error": str(e)}

    workload_name = {
        "fd": "FD RK3 step (cds=2), neutral closed channel, Re_tau=180, 256x256x257, zstretch=1, wall_rows=1",
        "fd_t": "FD RK3 step (cds=2) with temperature, stratified open channel, Re_tau=395, got=0.1, 256x256x257, zstretch=1, wall_rows=1",
        "sp": "spectral RK3 step (cds=3, skew-symmetric, 2/3 rule) with temperature, stratified open channel, Re_tau=395, got=0.1, top dT/dz=1, 256x256x257, zstretch=1, wall_rows=1",
        "poisson": "standalone pressure Poisson solve, 256x256x257, zstretch=1",
    }[args.workload]
    if args.zorder != 2:
        workload_name += f", {'6th' if args.zorder == 6 else '4th'}-order z operators (zorder={args.zorder})"
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if mixed_main else "f64",
        "data": "synthetic (seeded log-law + noise state; no checkpoint)",
        "config": {"workload": workload_name, "nx": prm.nx, "ny": prm.ny, "nz": prm.nz,
                   "cells": n, "precision": args.precision, "zorder": args.zorder},
        "roofline": roofline, "step_roofline": step_roofline, "cpu_baseline": cpu,
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
        "state_finite": finite, "div": div,
    }
    if getattr(ch, "tuning", None):
        out["config"]["launch_shapes"] = ch.tuning
    if other:
        out["mixed" if not mixed_main else "fp64"] = other
    if cpu and "value" in cpu and cpu["value"]:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    ch.close()
    if args.workload == "fd" and not args.no_extras and args.zorder == 2 and not mixed_main:
        # the metric's third part: the adjoint/forward cost ratio (128x128x129,
        # 400 steps, checkpoint every 20), and the 6th-order-z spectral step
        a = adjoint_case(400, 20)
        for key in ("K", "cells", "t_grad"):
            a.pop(key)
        out["adjoint"] = a
        c3 = config3_case(5, 1, 0)
        f64 = c3.pop("_fields")
        out["config3"] = c3
        try:
            c3m = config3_case(5, 1, 1)
            fm = c3m.pop("_fields")
            c3m["rel_l2_vs_fp64"] = {k: float(np.linalg.norm(fm[k] - f64[k]) / np.linalg.norm(f64[k]))
                                     for k in f64}
            c3m["wall_shear_rel_err"] = abs(c3m["wall_shear_bottom"] - c3["wall_shear_bottom"]) / abs(
                c3["wall_shear_bottom"])
            c3m["speedup_vs_fp64"] = round(c3["ms_per_step"] / c3m["ms_per_step"], 3)
            out["config3_mixed"] = c3m
        except Exception as e:  # the mixed leg never gates the line
            out["config3_mixed"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
