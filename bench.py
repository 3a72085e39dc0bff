#!/usr/bin/env python3
"""Benchmark of the per-step hot path on one MI355X: SSPRK(3,3) steps of the forced turbulent
channel at 4096x2049 FP64 (the reference solver's operator sequence, main.cu:434-541, with the
adaptive dt chosen on the device).

    python bench.py [--gpus 1] [--steps 500] [--warmup 200]

Prints one JSON line: throughput in grid-point steps per second, the roofline of the dominant
kernel (HIP events on the launch stream during the timed steps, algorithmic bytes per launch) and
the CPU baseline (the FP64 OpenMP oracle timed on a bounded sample of the same workload on this
host's cores), and under "les" the learned-closure LES leg (512x257 against the 4096x2049 DNS,
about 3 minutes, run in a child process; --no-les leaves it out).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nse2d-channel-gpu_amd"))
import nse2d_hip as nh  # noqa: E402

METRIC = "grid-point·steps/s at 4096×2049 FP64 (% HBM roofline); LES speedup at ≤2% Re_τ err"
UNIT = "grid-point·steps/s"
HBM_PEAK_GBS = 8000.0

# Forced turbulent channel, 4096 x 2049 FP64: Lx = 2 pi, h = 1, tanh beta = 2, Ub = 1, Re_b = 2e4,
# Kolmogorov forcing F0 = 0.05 at n = 12, drag 1e-3, laminar profile + band-limited (kx <= 32)
# seeded noise of amplitude 0.05 tapered by (1-y^2)^2, adaptive dt (CFL 0.5, viscous limit).
CFG = dict(Nx=4096, Ny=2049, Lx=2 * math.pi, h=1.0, beta=2.0, Ub=1.0, Reb=2e4, F0=0.05, nforce=12, drag=1e-3,
           amp=0.05, band_kmax=32, seed=20241203, cfl=0.5, cvisc=0.5, dt_max=1e-2)

# Algorithmic HBM bytes per grid point for one launch of each kernel class (compulsory traffic,
# FP64): stage kernel reads psi, omega_s (+ omega^n in stages 2 and 3) and writes omega_{s+1}
# (24/32/32 B, mean 88/3); forward FFT reads the real row (8 B) and writes the half spectrum
# (8 B); Thomas reads and writes the half spectrum (16 B); inverse FFT 16 B; CFL reads psi (8 B).
ALG_BYTES = {"stage": 88.0 / 3.0, "fft_forward": 16.0, "thomas": 16.0, "fft_inverse": 16.0, "cfl": 8.0}


def ic_args(c):
    return ["--Nx", str(c["Nx"]), "--Ny", str(c["Ny"]), "--Lx", repr(c["Lx"]), "--h", repr(c["h"]),
            "--Ub", repr(c["Ub"]), "--Reb", repr(c["Reb"]), "--init", "lamband", "--amp", repr(c["amp"]),
            "--band_kmax", str(c["band_kmax"]), "--seed", str(c["seed"]), "--stretch", "tanh",
            "--beta", repr(c["beta"]), "--F0", repr(c["F0"]), "--nforce", str(c["nforce"]),
            "--drag", repr(c["drag"]), "--cfl", repr(c["cfl"]), "--cvisc", repr(c["cvisc"])]


def cargv(args):
    return (ctypes.c_char_p * len(args))(*[a.encode() for a in args])


def setup_channel(c):
    """Metrics and IC through the library's host code, state on the device."""
    L = nh.lib()
    Nx, Ny = c["Nx"], c["Ny"]
    m = {k: np.zeros(Ny if k in ("y", "a_node", "w_node") else Ny - 1) for k in
         ("y", "a_node", "a_edge", "dy_edge", "w_node")}
    rc = L.nse2d_host_metrics(Ny, c["h"], b"tanh", c["beta"], *[m[k].ctypes.data for k in
                              ("y", "a_node", "a_edge", "dy_edge", "w_node")])
    if rc:
        raise RuntimeError(nh.last_error())
    w0 = np.zeros(Nx * Ny)
    args = ic_args(c)
    if L.nse2d_host_init_omega(len(args), cargv(args), w0.ctypes.data):
        raise RuntimeError(nh.last_error())
    dm = {k: nh.DeviceBuffer.from_array(v) for k, v in m.items()}
    d = nh.ChannelDesc()
    d.Nx, d.Ny, d.Lx, d.h = Nx, Ny, c["Lx"], c["h"]
    d.nu = 2.0 * c["h"] * c["Ub"] / c["Reb"]
    d.F0, d.lin_drag, d.nforce = c["F0"], c["drag"], c["nforce"]
    d.psi_bot, d.psi_top = 0.0, 2.0 * c["h"] * c["Ub"]
    d.cfl, d.cvisc, d.dt_max = c["cfl"], c["cvisc"], c["dt_max"]
    d.dy_min = float(np.min(np.abs(m["dy_edge"])))
    d.d_a_node, d.d_a_edge, d.d_y_node = dm["a_node"].ptr, dm["a_edge"].ptr, dm["y"].ptr
    d.d_dy_edge, d.d_w_node = dm["dy_edge"].ptr, dm["w_node"].ptr
    ch = L.nse2d_channel_create(ctypes.byref(d), None)
    if not ch:
        raise RuntimeError(nh.last_error())
    nh.upload_ptr(L.nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
    L.nse2d_channel_init(ch)
    return ch, dm


def read_profile(ch):
    ms = (ctypes.c_double * 6)()
    n = (ctypes.c_longlong * 6)()
    nh.lib().nse2d_channel_profile_read(ch, ms, n)
    return {nh.PROF_CLASSES[k]: (ms[k], n[k]) for k in range(6)}


def cpu_model():
    """The host CPU's model name (the lscpu "Model name" field, read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(c, budget_s):
    """The oracle (FP64 OpenMP restatement) on this host: one warm step, then as many steps as fit
    in about budget_s seconds (at least 1)."""
    lib_path = os.path.join(ROOT, "oracle", "liboracle_nse2d.so")
    if not os.path.isfile(lib_path):
        return None
    L = ctypes.CDLL(lib_path)
    L.orc_time_steps.restype = ctypes.c_double
    L.orc_time_steps.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int]
    L.orc_num_threads.restype = ctypes.c_int
    args = ic_args(c)
    t1 = L.orc_time_steps(len(args), cargv(args), 1, 1)  # includes one warm step
    steps = int(max(1, min(1000, budget_s / max(t1, 1e-3))))
    el = L.orc_time_steps(len(args), cargv(args), 1, steps)
    pts = c["Nx"] * c["Ny"] * steps
    return {"value": pts / el, "unit": UNIT, "cores": int(L.orc_num_threads()), "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{steps} SSPRK(3,3) steps (adaptive dt) of the {c['Nx']}x{c['Ny']} workload after one "
                      f"warm step, FP64 OpenMP oracle, {el:.2f} s"}


# Batched ensemble: 64 independent channels of uneven size (32 at 512x257, 24 at 1024x513, 8 at
# 2048x1025), tanh beta = 2, Re_b log-uniform in [2e3, 1e4] (seeded), own IC seed each, forcing and
# drag as the single channel, adaptive dt per member.
ENSEMBLE = [(512, 257, 32), (1024, 513, 24), (2048, 1025, 8)]


def ensemble_members(seed=20241203):
    rng = np.random.default_rng(seed)
    out = []
    for Nx, Ny, count in ENSEMBLE:
        for _ in range(count):
            out.append((Nx, Ny, float(np.exp(rng.uniform(np.log(2e3), np.log(1e4)))), 1000 + len(out)))
    return out


def run_ensemble(a):
    """One line of JSON: aggregate grid-point steps/s of the ensemble (autotuned, graph-replayed
    steps) and of the same members advanced one after another as single channels."""
    L = nh.lib()
    members = ensemble_members()
    grids, keep, descs, ics = {}, [], [], []
    for Nx, Ny, Reb, seed in members:
        c = dict(CFG, Nx=Nx, Ny=Ny, Reb=Reb, seed=seed)
        if (Nx, Ny) not in grids:
            m = {k: np.zeros(Ny if k in ("y", "a_node", "w_node") else Ny - 1) for k in
                 ("y", "a_node", "a_edge", "dy_edge", "w_node")}
            if L.nse2d_host_metrics(Ny, 1.0, b"tanh", c["beta"], *[m[k].ctypes.data for k in
                                    ("y", "a_node", "a_edge", "dy_edge", "w_node")]):
                raise RuntimeError(nh.last_error())
            grids[(Nx, Ny)] = (m, {k: nh.DeviceBuffer.from_array(v) for k, v in m.items()})
        m, dm = grids[(Nx, Ny)]
        w0 = np.zeros(Nx * Ny)
        args = ic_args(c)
        if L.nse2d_host_init_omega(len(args), cargv(args), w0.ctypes.data):
            raise RuntimeError(nh.last_error())
        d = nh.ChannelDesc()
        d.Nx, d.Ny, d.Lx, d.h = Nx, Ny, c["Lx"], c["h"]
        d.nu, d.F0, d.lin_drag, d.nforce = 2.0 * c["h"] * c["Ub"] / Reb, c["F0"], c["drag"], c["nforce"]
        d.psi_bot, d.psi_top = 0.0, 2.0 * c["h"] * c["Ub"]
        d.cfl, d.cvisc, d.dt_max = c["cfl"], c["cvisc"], c["dt_max"]
        d.dy_min = float(np.min(np.abs(m["dy_edge"])))
        d.d_a_node, d.d_a_edge, d.d_y_node = dm["a_node"].ptr, dm["a_edge"].ptr, dm["y"].ptr
        d.d_dy_edge, d.d_w_node = dm["dy_edge"].ptr, dm["w_node"].ptr
        descs.append(d)
        ics.append(w0)
    pts = sum(Nx * Ny for Nx, Ny, _, _ in members)
    # the members one after another as single channels (same shapes as the untuned ensemble)
    serial_s, serial_steps = 0.0, max(10, a.steps // 10)
    for d, w0 in zip(descs, ics):
        ch = L.nse2d_channel_create(ctypes.byref(d), None)
        nh.upload_ptr(L.nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
        L.nse2d_channel_init(ch)
        L.nse2d_channel_step_auto(ch, 5)
        L.nse2d_device_sync()
        t0 = time.perf_counter()
        L.nse2d_channel_step_auto(ch, serial_steps)
        L.nse2d_device_sync()
        serial_s += (time.perf_counter() - t0) / serial_steps
        L.nse2d_channel_destroy(ch)
    arr = (nh.ChannelDesc * len(descs))(*descs)
    e = L.nse2d_ensemble_create(arr, len(descs), None)
    if not e:
        raise RuntimeError(nh.last_error())
    for k, w0 in enumerate(ics):
        nh.upload_ptr(L.nse2d_ensemble_field(e, k, nh.FIELD_OMEGA), w0)
    L.nse2d_ensemble_init(e)
    rep = ctypes.create_string_buffer(1024)
    if a.no_autotune:
        rep.value = b"off (shapes from NSE2D_STAGE_TILE / NSE2D_FFT_RADIX / NSE2D_THOMAS_KC or the defaults)"
    else:
        L.nse2d_ensemble_autotune(e, 3, rep, 1024)
    L.nse2d_ensemble_step_auto(e, a.warmup)
    L.nse2d_ensemble_sync(e)
    t0 = time.perf_counter()
    L.nse2d_ensemble_step_auto(e, a.steps)
    L.nse2d_ensemble_sync(e)
    el = time.perf_counter() - t0
    sc = nh.download_ptr(L.nse2d_ensemble_scalars(e), 8 * len(descs)).reshape(-1, 8)
    if not np.all(np.isfinite(sc[:, :2])):
        raise RuntimeError("non-finite member time or dt")
    L.nse2d_ensemble_destroy(e)
    value = pts * a.steps / el
    serial = pts / serial_s
    out = {
        "metric": "aggregate grid-point·steps/s, batched ensemble of 64 channels FP64", "value": value,
        "unit": UNIT, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "ensemble: 32 x 512x257 + 24 x 1024x513 + 8 x 2048x1025, tanh beta=2, Re_b "
                               "log-uniform [2e3, 1e4], own seed and adaptive dt per member",
                   "members": len(descs), "grid_points": pts},
        "serial_single_channel_value": serial, "speedup_vs_serial": value / serial,
        "step_roofline_frac": 240.0 * value / (HBM_PEAK_GBS * 1e9),
        "autotune": rep.value.decode(), "member_t_range": [float(sc[:, 0].min()), float(sc[:, 0].max())],
    }
    print(json.dumps(out), flush=True)


LES_CONFIG = ("learned-closure LES 512x257 (tanh beta=2, Re_b=2e4, F0=0.05 n=12, drag 1e-3) trained on "
              "4096x2049 DNS filtered 8x in x (sharp spectral) and 8x in mapped y (box)")


def les_record(r, a):
    """The LES leg's figures: speedup per unit simulated time of the coarse run with the closure
    evaluated once per step, and for each coarse run the Re_tau, U(y), budget and trajectory
    errors against the filtered DNS, over the training window and over the whole comparison."""
    runs = {}
    for name in ("closure_per_step", "closure_mean_free", "no_closure"):
        x = r["runs"][name]
        w = x.get("window", {})
        runs[name] = {
            "finite": x["finite"], "step_budget_exceeded": x.get("step_budget_exceeded", False),
            "re_tau_err_window": w.get("re_tau_rel_err"), "U_err_window": w.get("U_rel_err"),
            "budget_err_window": w.get("budget"), "traj_err_window_end": x.get("traj_err_window_end"),
            "re_tau_err": x.get("re_tau_rel_err"), "U_err": x.get("U_rel_err"), "budget_err": x.get("budget"),
            "re_tau_err_thirds": x.get("re_tau_rel_err_thirds"), "traj_err_end": x.get("traj_err_end"),
            "wall_s_per_unit_time": x["wall_per_time"],
            "within_2pct": bool(x["finite"] and (x.get("re_tau_rel_err") or 1.0) <= 0.02
                                and (x.get("U_rel_err") or 1.0) <= 0.02)}
    t0 = a.les_spinup
    dns = r["dns_wall_per_time"]
    for x in runs.values():
        x["speedup"] = dns / x["wall_s_per_unit_time"] if x["wall_s_per_unit_time"] else None
    # the metric's figure: the fastest closure run whose Re_tau and U(y) stay within 2 % of the
    # filtered DNS over the whole comparison window (null when none does)
    ok = [n for n in ("closure_per_step", "closure_mean_free") if runs[n]["within_2pct"] and runs[n]["speedup"]]
    best = max(ok, key=lambda n: runs[n]["speedup"]) if ok else None
    return {
        "metric": "LES wall-clock speedup over DNS per unit simulated time at <=2% Re_tau and U(y) error "
                  "(512x257 closure LES vs 4096x2049 DNS)",
        "speedup": runs[best]["speedup"] if best else None, "unit": "x", "speedup_run": best,
        "speedup_closure_per_step_unconstrained": r["speedup_vs_dns_per_step_closure"],
        "speedup_closure_every_stage": r["speedup_vs_dns"],
        "speedup_no_closure": runs["no_closure"]["speedup"],
        "dns_wall_s_per_unit_time": dns,
        "config": {"workload": LES_CONFIG, "pairs": a.les_pairs, "batch": 32, "train_steps": a.les_train_steps,
                   "training_window": [t0, t0 + a.les_window], "comparison_window": [t0, t0 + a.les_compare],
                   "closure": "evaluated once per step, held over the three stages"},
        "re_tau_dns_filtered": r["re_tau_dns_filtered"], "re_tau_dns_fine": r["re_tau_dns_fine"],
        "holdout_mse_cnn": r["holdout_mse_cnn"], "holdout_mse_zero": r["holdout_mse_zero"],
        "runs": runs,
    }


def run_les(a):
    """One line of JSON: the learned-closure LES at 512x257 against the 4096x2049 DNS it is trained
    on (DNS -> 2000 filtered pairs -> Adam, batch 32 -> coarse runs with and without the closure
    from the filtered DNS state): wall-clock speedup per unit of simulated time, Re_tau, U(y),
    budget and trajectory errors against the filtered DNS."""
    import les_experiment as lx
    fine = dict(CFG)
    coarse = dict(CFG, Nx=512, Ny=257)
    r = lx.run(fine, coarse, spinup_time=a.les_spinup, window_time=a.les_window, n_pairs=a.les_pairs,
               holdout=max(1, a.les_pairs // 20), train_steps=a.les_train_steps, batch=32, lr=1e-3,
               n_samples=min(100, a.les_pairs), log=lambda m: print("les: " + m, file=sys.stderr, flush=True),
               compare_time=a.les_compare)
    rec = les_record(r, a)
    out = {"metric": rec["metric"], "value": rec["speedup"], "unit": "x", "n_gpus": 1, "higher_is_better": True,
           "vs_baseline": None, "dtype": "f64 solver, bf16/fp32 CNN", "data": "synthetic (DNS generated in the run)",
           "config": rec["config"], "les": rec, "details": r}
    print(json.dumps(out), flush=True)


def les_leg(a):
    """The LES leg of the default line, run as a child process (its own GPU context, a time limit)
    so that whatever happens there the channel figures are still printed: the child's record, or
    {"error": ...}."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", "les",
           "--les-spinup", repr(a.les_spinup), "--les-window", repr(a.les_window),
           "--les-compare", repr(a.les_compare), "--les-pairs", str(a.les_pairs),
           "--les-train-steps", str(a.les_train_steps)]
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, timeout=a.les_timeout, text=True)
    except subprocess.TimeoutExpired:
        return {"error": f"LES leg exceeded {a.les_timeout:.0f} s"}
    if p.returncode != 0:
        return {"error": f"LES leg exited with status {p.returncode}"}
    try:
        return json.loads(p.stdout.strip().splitlines()[-1])["les"]
    except (ValueError, KeyError, IndexError) as e:
        return {"error": f"LES leg printed no record ({e})"}


def run_cnn(a):
    """One line of JSON: the closure network at the LES grid (257x512): inference of one image (the
    closure term of one coarse state) and one Adam training step at batch 32, with the bf16 MFMA
    roofline of the convolutions (dense peak 2.5 PFLOP/s)."""
    L = nh.lib()
    H, W, B = 257, 512, 32
    c = L.nse2d_cnn_create(H, W, B, 1, None)
    if not c:
        raise RuntimeError(nh.last_error())
    rng = np.random.default_rng(0)
    x = nh.DeviceArray(rng.normal(size=(B, H, W, 4)).astype(np.float32))
    y = nh.DeviceArray(rng.normal(size=(B, H, W)).astype(np.float32))
    o = nh.DeviceArray(shape=(B, H, W), dtype=np.float32)
    # executed convolution FLOPs per pixel: first layer 32 (padded) -> 64, four 64 -> 64, last 64 -> 1
    flop_px = 2 * 9 * (32 * 64 + 4 * 64 * 64 + 64)
    for _ in range(a.warmup):
        L.nse2d_cnn_infer(c, x.ptr, o.ptr, 1)
    L.nse2d_device_sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        L.nse2d_cnn_infer(c, x.ptr, o.ptr, 1)
    L.nse2d_device_sync()
    t_inf = (time.perf_counter() - t0) / a.steps
    L.nse2d_cnn_train_step(c, x.ptr, y.ptr, B, 1e-4)
    n_train = max(3, a.steps // 10)
    t0 = time.perf_counter()
    for _ in range(n_train):
        L.nse2d_cnn_train_step(c, x.ptr, y.ptr, B, 1e-4)
    t_train = (time.perf_counter() - t0) / n_train
    L.nse2d_cnn_destroy(c)
    inf_tflops = H * W * flop_px / t_inf / 1e12
    train_tflops = 3 * B * H * W * flop_px / t_train / 1e12
    out = {"metric": "closure network inference time per coarse state (257x512, batch 1)", "value": t_inf * 1e3,
           "unit": "ms", "higher_is_better": False, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
           "config": {"workload": "6-layer 3x3 CNN, 64 channels, GELU, 257x512 images", "batch_train": B},
           "inference_tflops": inf_tflops, "train_step_ms": t_train * 1e3, "train_tflops": train_tflops,
           "roofline": {"bound": "mfma", "achieved": train_tflops, "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": train_tflops / 2500.0, "kernel": "training step, batch 32"}}
    print(json.dumps(out), flush=True)


def run_diag(a):
    """One line of JSON: one diagnostics row (the reductions of main.cu:565-692) at 4096x2049 on the
    channel's state after the warm-up, fused (nse2d_diag_row: one pass over psi, omega and rhs)
    against the separate launchers, with the fused row's HBM roofline (algorithmic bytes: psi,
    omega and rhs read once, 24 B per point)."""
    L = nh.lib()
    c = CFG
    Nx, Ny = c["Nx"], c["Ny"]
    ch, dm = setup_channel(c)
    L.nse2d_channel_step_auto(ch, a.warmup)
    L.nse2d_channel_step(ch, 0.0, 1e-5, nh.STEP_STORE_RHS)
    psi, w, rhs = (L.nse2d_channel_field(ch, k) for k in (nh.FIELD_PSI, nh.FIELD_OMEGA, nh.FIELD_RHS))
    inv_dx, inv_deta = Nx / c["Lx"], (Ny - 1) / 2.0
    out = nh.DeviceBuffer(32)
    sep = nh.DeviceBuffer(32)
    s = L.nse2d_stream_create()
    an, ae, yn, wn = dm["a_node"].ptr, dm["a_edge"].ptr, dm["y"].ptr, dm["w_node"].ptr

    def fused():
        if L.nse2d_diag_row(psi, w, rhs, Nx, Ny, inv_dx, inv_deta, c["h"], c["F0"], c["nforce"], an, ae, yn, wn,
                            out.ptr, None, s):
            raise RuntimeError(nh.last_error())

    def separate():
        b = sep.ptr
        L.launch_max_uv_from_psi_mapped(psi, Nx, Ny, inv_dx, inv_deta, an, b, b + 8, b + 16, s)
        L.launch_kinetic_energy_mapped(psi, Nx, Ny, inv_dx, inv_deta, an, wn, b + 24, s)
        L.launch_enstrophy_diss_mapped_full(w, Nx, Ny, inv_dx, inv_deta, an, wn, b + 32, b + 40, s)
        L.launch_omega_rhs_balance_mapped(w, rhs, Nx, Ny, wn, b + 48, s)
        L.launch_power_input_mapped(psi, Nx, Ny, inv_deta, c["h"], c["F0"], c["nforce"], an, yn, wn, b + 56, s)
        L.launch_energy_dissipation_mapped(psi, w, Nx, Ny, inv_dx, inv_deta, an, ae, wn, b + 64, s)
        L.launch_enstrophy_input_mapped_full(w, Nx, Ny, c["h"], c["F0"], c["nforce"], yn, wn, b + 72, s)
        L.launch_mean_wall_omega_omegay_mapped(w, Nx, Ny, inv_deta, an, b + 80, b + 88, s)
        L.launch_mean_wall_omega(w, Nx, Ny, b + 96, b + 104, s)

    res = {}
    for name, fn in (("fused", fused), ("separate", separate)):
        for _ in range(5):
            fn()
        L.nse2d_stream_sync(s)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            fn()
        L.nse2d_stream_sync(s)
        res[name] = (time.perf_counter() - t0) / a.steps
    same = bool(np.array_equal(out.download()[:14], sep.download()[:14]))
    L.nse2d_stream_destroy(s)
    L.nse2d_channel_destroy(ch)
    alg = 24.0 * Nx * Ny
    achieved = alg / res["fused"] / 1e9
    print(json.dumps({
        "metric": "diagnostics row time at 4096x2049 FP64 (fused one-pass reductions)", "value": res["fused"] * 1e6,
        "unit": "us", "higher_is_better": False, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (the channel's state after the warm-up)",
        "config": {"workload": "one diagnostics row of the 4096x2049 forced channel: CFL maxima, K, Omega, eta, "
                               "<omega rhs>, Pin, epsilon, P_Omega, wall means"},
        "separate_launchers_us": res["separate"] * 1e6, "speedup": res["separate"] / res["fused"],
        "fused_equals_separate_bitwise": same,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_row": alg,
                     "note": "host-timed per row (two launches), so launch gaps are included"}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["channel", "ensemble", "les", "cnn", "diag"], default="channel",
                    help="channel: the 4096x2049 forced channel (default); ensemble: 64 independent channels; "
                         "les: learned-closure LES at 512x257 trained on the 4096x2049 DNS; cnn: the closure "
                         "network alone; diag: one fused diagnostics row")
    ap.add_argument("--les-spinup", type=float, default=4.0)
    ap.add_argument("--les-window", type=float, default=4.0)
    ap.add_argument("--les-compare", type=float, default=12.0,
                    help="LES: length of the comparison window from the start of the training window")
    ap.add_argument("--les-timeout", type=float, default=600.0, help="channel: time limit of the LES leg")
    ap.add_argument("--no-les", action="store_true", help="channel: leave out the LES leg")
    ap.add_argument("--les-pairs", type=int, default=2000)
    ap.add_argument("--les-train-steps", type=int, default=1250)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--time-every", type=int, default=10,
                    help="in the timed region, bracket the dominant kernel with events on every k-th step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="ensemble: keep the configured kernel shapes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run 5 device-to-device copies of 256 MiB (a plain copy kernel of known size, "
                         "for scaling HBM counter readings)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (the channel "
              "would shard by y-slabs with a partitioned Thomas solve, which is not built)", file=sys.stderr)
        sys.exit(2)
    nh.require_gpu()
    L = nh.lib()
    if a.copy_calibration:
        n = 256 << 20
        src, dst = nh.DeviceBuffer(n // 8, init=np.ones(n // 8)), nh.DeviceBuffer(n // 8)
        for _ in range(5):
            L.nse2d_memcpy_d2d(dst.ptr, src.ptr, n)
        L.nse2d_device_sync()
        src.free()
        dst.free()
    if a.workload == "ensemble":
        run_ensemble(a)
        return
    if a.workload == "les":
        run_les(a)
        return
    if a.workload == "cnn":
        run_cnn(a)
        return
    if a.workload == "diag":
        run_diag(a)
        return
    c = CFG
    Nx, Ny = c["Nx"], c["Ny"]
    ch, keep = setup_channel(c)
    # warm-up, then a short fully profiled segment to find the dominant kernel class
    L.nse2d_channel_step_auto(ch, a.warmup)
    L.nse2d_channel_profile(ch, (1 << 6) - 1)
    L.nse2d_channel_step_auto(ch, 10)
    prof = read_profile(ch)
    per_step = {k: v[0] / 10.0 for k, v in prof.items()}
    # the production stages run fused with the forward row transform when the plan allows it:
    # then the stage class also writes the half spectrum and no forward launches are seen
    fused = prof["fft_forward"][1] == 0
    alg = dict(ALG_BYTES)
    if fused:
        alg["stage"] += 8.0
    dom = max(alg, key=lambda k: prof[k][0])
    dom_bit = nh.PROF_CLASSES.index(dom)
    # timed region: only the dominant class is bracketed with events, on every k-th step (the
    # events themselves add time to the steps they bracket)
    L.nse2d_channel_profile(ch, 1 << dom_bit)
    L.nse2d_channel_profile_every(ch, a.time_every)
    L.nse2d_device_sync()
    t0 = time.perf_counter()
    L.nse2d_channel_step_auto(ch, a.steps)
    L.nse2d_device_sync()
    t1 = time.perf_counter()
    dom_ms, dom_n = read_profile(ch)[dom]
    L.nse2d_channel_profile(ch, 0)
    L.nse2d_channel_profile_every(ch, 1)
    w = nh.download_ptr(L.nse2d_channel_field(ch, nh.FIELD_OMEGA), Nx * Ny)
    if not np.all(np.isfinite(w)):
        raise RuntimeError("non-finite vorticity after the timed steps")
    t_sim = nh.download_ptr(L.nse2d_channel_time_ptr(ch), 1)[0]
    elapsed = t1 - t0
    value = Nx * Ny * a.steps / elapsed
    avg_s = dom_ms / max(dom_n, 1) * 1e-3
    bytes_per_launch = alg[dom] * Nx * Ny
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.isfile(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get(dom)
    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "forced turbulent channel 4096x2049 FP64 (tanh beta=2, Re_b=2e4, F0=0.05 n=12, "
                               "drag 1e-3, band-limited kx<=32 noise IC, adaptive dt CFL 0.5)",
                   "Nx": Nx, "Ny": Ny, "sim_time_reached": t_sim},
        "roofline": {"bound": "hbm", "kernel": "stage+fft_forward" if (fused and dom == "stage") else dom,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                                       "this workload, scaled by a 256 MiB copy (not measured in this run)"
                                       if traffic is not None else None,
                     "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": bytes_per_launch,
                     "timed_launches": int(dom_n), "timed_every_k_steps": a.time_every},
        # 240 B/point: the unfused step (three stages, three Poisson solves, the CFL read); the
        # fused stage + transform skips one 8-byte read per point and stage (216 B/point moved)
        "step_alg_bytes_per_point": 240,
        "step_roofline_frac": 240.0 * value / (HBM_PEAK_GBS * 1e9),
        "step_bytes_moved_per_point": 216 if fused else 240,
        "step_moved_frac": (216 if fused else 240) * value / (HBM_PEAK_GBS * 1e9),
        "kernels_ms_per_step": per_step,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(c, a.cpu_seconds)
    L.nse2d_channel_destroy(ch)
    for buf in keep.values():
        buf.free()
    if not a.no_les:
        out["les"] = les_leg(a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
