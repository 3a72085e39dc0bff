"""Verification harness: the 21 cases and six checks of the reference's tests/run_tests.py,
restated for this project (run_tests.py:48-111 cases, :151-207 readers, :213-333 metrics,
:342-345 tolerances). A "solver" is any callable taking the reference command line (a list of
strings without the program name) and returning an exit code: the HIP driver (nse2d_main), the
nse2d executables, or the oracle.

Snapshots are this project's .bin + .json pair (the reference writes HDF5).
"""
import glob
import json
import math
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "reference")

LX = 4.0 * math.pi  # run_tests.py:48 (the solver's own default is 12.566370612)
H = 1.0
EIG_NU, EIG_T, EIG_NY, EIG_BETA, EIG_DT = 1e-2, 2.0, [33, 65, 129], [0, 2], 1e-4
TIM_NU, TIM_T, TIM_M, TIM_AU = 1e-4, 10.24, 8, 0.1
TIM_DT, TIM_DTREF, TIM_NY = [0.64, 0.32, 0.16, 0.08], 0.01, 129
POI_REB, POI_UB, POI_T = 500.0, 1.0, 5.0
MMS_NU, MMS_T, MMS_A0, MMS_EPS, MMS_OM, MMS_KX = 1e-2, 1.0, 1.0, 0.1, 1.0, 2
MMS_N, MMS_BETA = [64, 128, 256], [0, 2]
CON_DT, CON_T = [4e-3, 2e-3], 2.0
CH_REB, CH_F0, CH_NF, CH_DRAG, CH_T = 5000.0, 0.05, 12, 1e-3, 20.0
CH_DIAG, CH_TREF = 0.05, 2.0

REF_TOL = {"eigenmode_space": 0.02, "eigenmode_time": 0.05, "poiseuille": 0.05, "mms": 0.02,
           "conservation": 0.5, "channel": 1e-4}
REF_SKIP = {"channel": ("identity_max_over_eta_h", "eta_h_over_eta_late", "Phi_h_over_Phi_late")}
REF_ABS_FLOOR = 1e-13


def _stretch(beta):
    return ["--stretch", "none"] if beta == 0 else ["--stretch", "tanh", "--beta", str(beta)]


def mms_dt(n):
    return 5e-4 * (64.0 / n) ** 2


def cases():
    """(name, walls, args) of every case; walls is 'noslip' or 'freeslip'."""
    out = []
    for b in EIG_BETA:
        for ny in EIG_NY:
            out.append((f"eig_space_Ny{ny}_b{b}", "noslip",
                        ["--Nx", "64", "--Ny", str(ny), "--Ub", "0", "--nu", str(EIG_NU), "--init", "diffusion",
                         "--my", "1", "--amp", "1"] + _stretch(b) +
                        ["--no-adapt", "--dt", str(EIG_DT), "--tend", str(EIG_T), "--snap_save", str(EIG_T),
                         "--snap_fields", "omega_psi", "--diag_save", "0.5"]))
    for dt in TIM_DT + [TIM_DTREF]:
        name = "eig_time_ref" if dt == TIM_DTREF else f"eig_time_dt{dt:g}"
        out.append((name, "noslip",
                    ["--Nx", "64", "--Ny", str(TIM_NY), "--Ub", "0", "--nu", str(TIM_NU), "--init", "diffusion",
                     "--my", str(TIM_M), "--amp", str(TIM_AU), "--stretch", "none", "--no-adapt", "--dt", str(dt),
                     "--dtmax", "1", "--tend", str(TIM_T), "--snap_save", str(TIM_T), "--snap_fields", "omega",
                     "--diag_save", str(TIM_T)]))
    out.append(("poiseuille", "noslip",
                ["--Nx", "64", "--Ny", "129", "--Ub", str(POI_UB), "--Reb", str(POI_REB), "--init", "laminar"] +
                _stretch(2) + ["--tend", str(POI_T), "--diag_save", "0.5", "--snap_save", str(POI_T),
                               "--snap_fields", "omega_psi"]))
    for b in MMS_BETA:
        for n in MMS_N:
            out.append((f"mms_N{n}_b{b}", "noslip",
                        ["--Nx", str(n), "--Ny", str(n + 1), "--Ub", "0", "--nu", str(MMS_NU), "--init", "mms",
                         "--mmsA0", str(MMS_A0), "--mmsEps", str(MMS_EPS), "--mmsOm", str(MMS_OM),
                         "--mmsKx", str(MMS_KX)] + _stretch(b) +
                        ["--no-adapt", "--dt", repr(mms_dt(n)), "--tend", str(MMS_T), "--snap_save", str(MMS_T),
                         "--snap_fields", "omega_psi", "--diag_save", str(MMS_T)]))
    for dt in CON_DT:
        out.append((f"cons_dt{dt:g}", "freeslip",
                    ["--Nx", "256", "--Ny", "257", "--Ub", "0", "--nu", "0", "--init", "rand", "--amp", "0.5",
                     "--seed", "7", "--stretch", "none", "--no-adapt", "--dt", str(dt), "--tend", str(CON_T),
                     "--diag_save", "0.2", "--snap_save", "0.4", "--snap_fields", "omega_psi"]))
    out.append(("channel", "noslip",
                ["--Nx", "512", "--Ny", "257", "--Ub", "1", "--Reb", str(CH_REB), "--init", "mix", "--amp", "0.05",
                 "--F0", str(CH_F0), "--nforce", str(CH_NF), "--drag", str(CH_DRAG)] + _stretch(2) +
                ["--tend", str(CH_T), "--diag_save", str(CH_DIAG), "--snap_save", str(CH_T), "--snap_fields",
                 "omega_psi", "--closure_diag"]))
    return out


def case_args(name):
    for n, walls, a in cases():
        if n == name:
            return walls, a
    raise KeyError(name)


def solver_args(walls, args, outdir):
    extra = ["--walls", "freeslip"] if walls == "freeslip" else []
    return list(args) + extra + ["--outdir", outdir]


def run_case(solver, name, outroot, extra=()):
    walls, a = case_args(name)
    d = os.path.join(outroot, name)
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    rc = solver(solver_args(walls, list(a) + list(extra), d))
    if rc != 0:
        raise RuntimeError(f"case {name}: solver exit code {rc}")
    return d


# ---------------------------------------------------------------------------------------------
# readers
# ---------------------------------------------------------------------------------------------
def read_diag(d):
    a = np.genfromtxt(os.path.join(d, "diagnostics.csv"), delimiter=",", names=True)
    return np.atleast_1d(a)


def golden_diag(name):
    return read_diag(os.path.join(GOLDEN, "output", name))


def golden_profiles(name):
    return np.genfromtxt(os.path.join(GOLDEN, "output", name, "profiles_timeavg.csv"), delimiter=",", names=True)


def golden_reference():
    with open(os.path.join(GOLDEN, "reference.json")) as f:
        return json.load(f)


class Snap:
    pass


def read_snaps(d):
    """All snapshots of a run sorted by t; fields have shape [Ny, Nx]."""
    out = []
    for fn in sorted(glob.glob(os.path.join(d, "snaps", "snap_t*.json"))):
        with open(fn) as f:
            meta = json.load(f)
        s = Snap()
        s.t, s.Nx, s.Ny = float(meta["t"]), int(meta["Nx"]), int(meta["Ny"])
        s.Lx, s.h, s.nu = float(meta["Lx"]), float(meta["h"]), float(meta["nu"])
        raw = np.fromfile(fn[:-5] + ".bin", dtype="<f8")
        N = s.Nx * s.Ny
        fields = meta["fields"]
        for k, name in enumerate(fields):
            if name == "y":
                s.y = raw[k * N:k * N + s.Ny].copy()
            else:
                setattr(s, name, raw[k * N:(k + 1) * N].reshape(s.Ny, s.Nx))
        if not hasattr(s, "psi"):
            s.psi = None
        s.x = np.arange(s.Nx) * (s.Lx / s.Nx)
        out.append(s)
    if not out:
        raise FileNotFoundError(f"no snapshots in {os.path.join(d, 'snaps')}")
    return sorted(out, key=lambda s: s.t)


def last_snap(d, t_expected):
    s = read_snaps(d)[-1]
    if abs(s.t - t_expected) > 1e-6 * max(1.0, abs(t_expected)):
        raise RuntimeError(f"{d}: last snapshot at t={s.t}, expected {t_expected}")
    return s


def trap_weights(y):
    w = np.zeros_like(y)
    w[1:] += 0.5 * np.diff(y)
    w[:-1] += 0.5 * np.diff(y)
    return w


def norms(err, y):
    w = trap_weights(y)
    l2 = math.sqrt(float(np.sum(w[:, None] * err ** 2) / (err.shape[1] * np.sum(w))))
    return l2, float(np.max(np.abs(err)))


def order(e_coarse, e_fine, ratio=2.0):
    return math.log(e_coarse / e_fine) / math.log(ratio)


# ---------------------------------------------------------------------------------------------
# the six tests: each returns (metrics, [(description, ok, detail)])
# ---------------------------------------------------------------------------------------------
def t_eigenmode_space(out):
    m, checks = {}, []
    mu = math.pi / H
    for b in EIG_BETA:
        e2 = []
        for ny in EIG_NY:
            s = last_snap(os.path.join(out, f"eig_space_Ny{ny}_b{b}"), EIG_T)
            decay = math.exp(-EIG_NU * mu * mu * s.t)
            exact = -mu * np.cos(mu * (s.y + H)) * decay
            l2, li = norms(s.omega - exact[:, None], s.y)
            wall = 0.5 * (s.omega[0].mean() + s.omega[-1].mean())
            m[f"E2_Ny{ny}_b{b}"], m[f"Einf_Ny{ny}_b{b}"] = l2, li
            m[f"wall_rel_Ny{ny}_b{b}"] = abs(wall + mu * decay) / abs(mu * decay)
            e2.append(l2)
        p = [order(e2[i], e2[i + 1]) for i in range(len(e2) - 1)]
        checks.append((f"beta={b}: order in space {p}", 1.7 <= p[-1] <= 2.3, f"E2={e2}"))
    return m, checks


def t_eigenmode_time(out):
    m, checks = {}, []
    ref = last_snap(os.path.join(out, "eig_time_ref"), TIM_T)
    e2 = []
    for dt in TIM_DT:
        s = last_snap(os.path.join(out, f"eig_time_dt{dt:g}"), TIM_T)
        l2, _ = norms(s.omega - ref.omega, s.y)
        m[f"E2_dt{dt:g}"] = l2
        e2.append(l2)
    p = [order(e2[i], e2[i + 1]) for i in range(len(e2) - 1)]
    checks.append((f"order in time {p}", 2.5 <= 0.5 * (p[0] + p[1]) <= 3.5, f"E2={e2}"))
    return m, checks


def t_poiseuille(out):
    m, checks = {}, []
    r = read_diag(os.path.join(out, "poiseuille"))[-1]
    nu = 2.0 * H * POI_UB / POI_REB
    exact = {"tauw_bot": 3 * nu * POI_UB / H, "tauw_top": -3 * nu * POI_UB / H, "dpdx": -3 * nu * POI_UB / H ** 2,
             "epsilon": 3 * nu * POI_UB ** 2 / H ** 2, "K": 0.6 * POI_UB ** 2}
    for k, ex in exact.items():
        dev = abs(float(r[k]) - ex) / abs(ex)
        m[f"dev_{k}"] = dev
        checks.append((f"{k}={float(r[k]):+.6e} exact {ex:+.6e}", dev < 5e-3, f"dev {dev:.2e}"))
    bal = abs(-POI_UB * float(r["dpdx"]) - float(r["epsilon"])) / exact["epsilon"]
    m["energy_balance"] = bal
    checks.append(("energy balance -Ub dp/dx = epsilon", bal < 5e-3, f"mismatch {bal:.2e}"))
    return m, checks


def t_mms(out):
    m, checks = {}, []
    k = 2.0 * math.pi * MMS_KX / LX
    c = math.pi / (2.0 * H)
    for b in MMS_BETA:
        e2 = []
        for n in MMS_N:
            s = last_snap(os.path.join(out, f"mms_N{n}_b{b}"), MMS_T)
            amp = MMS_A0 * (1.0 + MMS_EPS * math.sin(MMS_OM * s.t))
            th = c * (s.y + H)
            g = k * k * np.sin(th) ** 2 - 2.0 * c * c * np.cos(2.0 * th)
            exact = amp * g[:, None] * np.sin(k * s.x)[None, :]
            l2, li = norms(s.omega - exact, s.y)
            m[f"E2_N{n}_b{b}"], m[f"Einf_N{n}_b{b}"] = l2, li
            e2.append(l2)
        p = [order(e2[i], e2[i + 1]) for i in range(len(e2) - 1)]
        checks.append((f"beta={b}: order {p}", 1.7 <= p[-1] <= 2.3, f"E2={e2}"))
    return m, checks


def t_conservation(out):
    m, checks = {}, []
    drift = {}
    for dt in CON_DT:
        snaps = read_snaps(os.path.join(out, f"cons_dt{dt:g}"))
        dy = 2.0 * H / (snaps[0].Ny - 1)
        Z = np.array([0.5 * dy * np.sum(s.omega ** 2) / s.Nx for s in snaps])
        E = np.array([0.5 * dy * np.sum(s.psi * s.omega) / s.Nx for s in snaps])
        dZ, dE = float(np.max(np.abs(Z / Z[0] - 1.0))), float(np.max(np.abs(E / E[0] - 1.0)))
        m[f"drift_enstrophy_dt{dt:g}"], m[f"drift_energy_dt{dt:g}"] = dZ, dE
        drift[dt] = (dE, dZ)
        checks.append((f"dt={dt:g}: drift energy {dE:.2e} enstrophy {dZ:.2e}", dE < 1e-6 and dZ < 1e-6,
                       f"{len(snaps)} snapshots"))
    a, b = drift[CON_DT[0]], drift[CON_DT[1]]
    checks.append(("drift decreases with dt", b[0] < a[0] and b[1] < a[1], f"{a} -> {b}"))
    return m, checks


CLOSURE_COLS = ["Omega_int_w", "dOmega_same_int_w", "A_w", "P_int_w", "eta_h", "Phi_h"]


def t_channel(out, t_end=CH_T):
    m, checks = {}, []
    d = read_diag(os.path.join(out, "channel"))
    missing = [c for c in CLOSURE_COLS if c not in d.dtype.names]
    if missing:
        raise RuntimeError(f"diagnostics.csv lacks {missing}: run with --closure_diag")
    r = d[np.isfinite(d["eta_h"])]
    resid = r["dOmega_same_int_w"] - (r["A_w"] - r["eta_h"] + r["Phi_h"] + r["P_int_w"] - 2.0 * CH_DRAG * r["Omega_int_w"])
    worst = float(np.max(np.abs(resid)) / np.mean(np.abs(r["eta_h"])))
    m["identity_max_over_eta_h"] = worst
    checks.append((f"enstrophy identity at {len(r)} rows", worst < 1e-6, f"max/mean eta_h = {worst:.2e}"))
    late = r[r["t"] >= 0.5 * t_end]
    if len(late):
        m["eta_h_over_eta_late"] = float(np.mean(late["eta_h"]) / np.mean(late["eta"]))
        m["Phi_h_over_Phi_late"] = float(np.mean(late["Phi_h"]) / np.mean(late["BOmega_wall"]))
    early = d[(d["t"] > 0) & (d["t"] <= CH_TREF + 1e-9)]
    for col in ("K", "Omega", "eta", "epsilon"):
        m[f"early_mean_{col}"] = float(np.mean(early[col]))
    m[f"K_at_t{CH_TREF:g}"] = float(early["K"][-1])
    m[f"Omega_at_t{CH_TREF:g}"] = float(early["Omega"][-1])
    fin = np.all([np.all(np.isfinite(d[c][1:])) for c in ("K", "Omega", "eta", "epsilon", "Re_tau")])
    checks.append((f"reached t={float(d['t'][-1]):g} with finite diagnostics",
                   bool(fin) and float(d["t"][-1]) > t_end - CH_DIAG - 1e-6, f"{len(d)} rows"))
    return m, checks


TESTS = [("eigenmode_space", t_eigenmode_space), ("eigenmode_time", t_eigenmode_time),
         ("poiseuille", t_poiseuille), ("mms", t_mms), ("conservation", t_conservation), ("channel", t_channel)]


def compare_reference(name, metrics, ref=None):
    """Mismatches of `metrics` against reference.json (run_tests.py:372-384)."""
    ref = ref if ref is not None else golden_reference()
    tol, bad = REF_TOL[name], []
    for k, v in metrics.items():
        if k in REF_SKIP.get(name, ()) or k not in ref.get(name, {}):
            continue
        r = ref[name][k]
        if max(abs(v), abs(r)) < REF_ABS_FLOOR:
            continue
        if abs(v - r) > tol * max(abs(r), REF_ABS_FLOOR):
            bad.append(f"{k}: {v:.6e} vs {r:.6e}")
    return bad


def check(name, outroot):
    """Run test `name` on the outputs under outroot: (passed, lines)."""
    fn = dict(TESTS)[name]
    metrics, checks = fn(outroot)
    lines = [f"[{'ok' if ok else 'FAILED'}] {desc} ({detail})" for desc, ok, detail in checks]
    bad = compare_reference(name, metrics)
    lines.append(f"[{'ok' if not bad else 'FAILED'}] {len(metrics)} values vs reference (tol {REF_TOL[name]:g})"
                 + ("" if not bad else ": " + "; ".join(bad)))
    return all(ok for _, ok, _ in checks) and not bad, lines, metrics


# ---------------------------------------------------------------------------------------------
# row-by-row comparison of diagnostics tables
# ---------------------------------------------------------------------------------------------
ROUNDOFF_COLS = {"utau_avg", "Re_tau", "Cf"}  # square roots of wall stress: noise when tau_w ~ 1e-16


def compare_rows(ref, got, rows=None, rel=5e-8, floor=1e-12):
    """Worst relative deviation per column between two diagnostics tables (first `rows` rows).
    Values below `floor` in both tables are round-off and skipped; utau/Re_tau/Cf are skipped
    where the wall stresses are below 1e-13 (pure round-off)."""
    n = min(len(ref), len(got)) if rows is None else rows
    bad = []
    for r in range(n):
        tiny_tau = abs(float(ref[r]["tauw_bot"])) < 1e-13 and abs(float(ref[r]["tauw_top"])) < 1e-13
        for c in ref.dtype.names:
            a, b = float(ref[r][c]), float(got[r][c])
            if math.isnan(a) and math.isnan(b):
                continue
            if abs(a) < floor and abs(b) < floor:
                continue
            if c in ROUNDOFF_COLS and tiny_tau:
                continue
            d = abs(a - b) / max(abs(a), 1e-300)
            if not d <= rel:
                bad.append((r, c, a, b, d))
    return bad
