"""The learned-closure LES experiment, driven through the C-ABI (libnse2d_hip.so):

  1. a DNS of the forced channel on the fine grid (nse2d_channel_*), advanced with the device dt;
  2. training pairs from its states: the filtered state and the closure target
     (nse2d_les_pair: box filter in mapped y, sharp spectral cut in x, FP64);
  3. Adam training of the CNN on the device (nse2d_les_train, bf16 MFMA kernels);
  4. coarse runs from the filtered DNS state, with the closure attached (nse2d_les_attach: in
     every SSPRK stage, or once per step) and without it, whose mean profile U(y), Re_tau and
     energy/enstrophy budget (K, Omega, the enstrophy dissipation eta and the forcing's enstrophy
     input P_Omega, the reductions of main.cu:627-629 through nse2d_diag_row) are compared with
     those of the filtered DNS over the training window and, optionally, a longer comparison
     window, and whose wall-clock time per unit of simulated time is compared with the DNS's.

No reference counterpart (the reference runs DNS only). Nothing here runs on the CPU beyond the
bookkeeping of a few profiles; there is no fallback.
"""
import ctypes
import math
import time

import numpy as np

import nse2d_hip as nh


def _cargv(args):
    return (ctypes.c_char_p * len(args))(*[a.encode() for a in args])


def ic_args(c):
    """Command-line style IC arguments of nse2d_host_init_omega (laminar profile plus band-limited
    seeded noise tapered by (1-y^2)^2)."""
    return ["--Nx", str(c["Nx"]), "--Ny", str(c["Ny"]), "--Lx", repr(c["Lx"]), "--h", repr(c["h"]),
            "--Ub", repr(c["Ub"]), "--Reb", repr(c["Reb"]), "--init", "lamband", "--amp", repr(c["amp"]),
            "--band_kmax", str(c["band_kmax"]), "--seed", str(c["seed"]), "--stretch", "tanh",
            "--beta", repr(c["beta"]), "--F0", repr(c["F0"]), "--nforce", str(c["nforce"]),
            "--drag", repr(c["drag"]), "--cfl", repr(c["cfl"]), "--cvisc", repr(c["cvisc"])]


class StepBudgetExceeded(RuntimeError):
    """A coarse run needed more steps than its budget (its dt collapsed)."""


class Channel:
    """A channel on the device: metrics (host and device), descriptor and handle."""

    def __init__(self, c, stream=None, ic=True):
        L = nh.lib()
        self.c = dict(c)
        Nx, Ny = c["Nx"], c["Ny"]
        self.Nx, self.Ny = Nx, Ny
        self.m = {k: np.zeros(Ny if k in ("y", "a_node", "w_node") else Ny - 1) for k in
                  ("y", "a_node", "a_edge", "dy_edge", "w_node")}
        if L.nse2d_host_metrics(Ny, c["h"], b"tanh", c["beta"], *[self.m[k].ctypes.data for k in
                                ("y", "a_node", "a_edge", "dy_edge", "w_node")]):
            raise RuntimeError(nh.last_error())
        self.dm = {k: nh.DeviceBuffer.from_array(v) for k, v in self.m.items()}
        d = nh.ChannelDesc()
        d.Nx, d.Ny, d.Lx, d.h = Nx, Ny, c["Lx"], c["h"]
        d.nu = 2.0 * c["h"] * c["Ub"] / c["Reb"]
        d.F0, d.lin_drag, d.nforce = c["F0"], c["drag"], c["nforce"]
        d.psi_bot, d.psi_top = 0.0, 2.0 * c["h"] * c["Ub"]
        d.cfl, d.cvisc, d.dt_max = c["cfl"], c["cvisc"], c.get("dt_max", 1e-2)
        d.dy_min = float(np.min(np.abs(self.m["dy_edge"])))
        d.d_a_node, d.d_a_edge, d.d_y_node = self.dm["a_node"].ptr, self.dm["a_edge"].ptr, self.dm["y"].ptr
        d.d_dy_edge, d.d_w_node = self.dm["dy_edge"].ptr, self.dm["w_node"].ptr
        self.desc = d
        self.nu = d.nu
        self.ch = L.nse2d_channel_create(ctypes.byref(d), stream)
        if not self.ch:
            raise RuntimeError(nh.last_error())
        self._su = nh.DeviceBuffer(5 * Ny)
        self._wall = nh.DeviceBuffer(2)
        if ic:
            w0 = np.zeros(Nx * Ny)
            args = ic_args(c)
            if L.nse2d_host_init_omega(len(args), _cargv(args), w0.ctypes.data):
                raise RuntimeError(nh.last_error())
            self.set_omega(w0)

    def field(self, which):
        return nh.lib().nse2d_channel_field(self.ch, which)

    def set_omega(self, w, t=0.0):
        """Upload omega, solve psi and close the walls (main.cu:211-213), set the time."""
        nh.upload_ptr(self.field(nh.FIELD_OMEGA), w)
        nh.lib().nse2d_channel_init(self.ch)
        nh.upload_ptr(nh.lib().nse2d_channel_time_ptr(self.ch), np.array([t]))

    def time(self):
        return float(nh.download_ptr(nh.lib().nse2d_channel_time_ptr(self.ch), 1)[0])

    def step(self, n):
        nh.lib().nse2d_channel_step_auto(self.ch, int(n))

    def mean_u(self, psi_ptr=None):
        """x-mean of u = a_j d psi / d eta (row sums of diag.cu:468-499, clamped walls) per row."""
        L, Nx, Ny = nh.lib(), self.Nx, self.Ny
        p = psi_ptr or self.field(nh.FIELD_PSI)
        b = self._su.ptr
        L.launch_row_moments_from_psi_mapped(p, Nx, Ny, Nx / self.c["Lx"], (Ny - 1) / 2.0, self.dm["a_node"].ptr,
                                             b, b + 8 * Ny, b + 16 * Ny, b + 24 * Ny, b + 32 * Ny, None)
        return self._su.download()[:Ny] / Nx

    def re_tau(self):
        """Re_tau = u_tau h / nu with u_tau = sqrt((|tau_b| + |tau_t|) / 2), tau = nu <omega_wall>
        (main.cu:664-675)."""
        L = nh.lib()
        b = self._wall.ptr
        L.launch_mean_wall_omega(self.field(nh.FIELD_OMEGA), self.Nx, self.Ny, b, b + 8, None)
        wb, wt = self._wall.download()
        utau = math.sqrt(0.5 * (abs(self.nu * wb) + abs(self.nu * wt)))
        return utau * self.c["h"] / self.nu

    def rule_dt(self):
        """dt of the reference's host rule (main.cu:440-458) for the current psi."""
        L, d = nh.lib(), self.desc
        b = self._su.ptr
        L.launch_max_uv_from_psi_mapped(self.field(nh.FIELD_PSI), self.Nx, self.Ny, self.Nx / d.Lx,
                                        (self.Ny - 1) / 2.0, self.dm["a_node"].ptr, b, b + 8, b + 16, None)
        umax, _, vmaxm = self._su.download()[:3]
        dx, deta = d.Lx / self.Nx, 2.0 / (self.Ny - 1)
        dt_x = d.cfl * dx / umax if umax > 0 else 1e30
        dt_y = d.cfl * deta / vmaxm if vmaxm > 0 else 1e30
        dt_v = d.cvisc * 0.5 / (d.nu * (1.0 / dx ** 2 + 1.0 / d.dy_min ** 2)) if d.nu > 0 else 1e30
        return min(d.dt_max, dt_x, dt_y, dt_v)

    # optional step budget of advance_to (a closure that destabilises the run drives dt down)
    max_steps = None
    nsteps = 0

    def advance_to(self, t, t_target):
        """Host-dt steps from t to exactly t_target (the last steps clipped as main.cu:466-480).
        Raises StepBudgetExceeded once more than max_steps steps were taken since nsteps was reset."""
        while t < t_target - 1e-12 * max(1.0, abs(t_target)):
            if self.max_steps is not None and self.nsteps >= self.max_steps:
                raise StepBudgetExceeded(f"{self.nsteps} steps by t = {t:.4f}")
            self.nsteps += 1
            dt = self.rule_dt()
            rem = t_target - t
            if rem <= dt:
                dt = rem
            elif rem < 2.0 * dt:
                dt = 0.5 * rem
            nh.lib().nse2d_channel_step(self.ch, t, dt, 0)
            t = t + dt if dt != rem else t_target
        return t

    def omega(self):
        return nh.download_ptr(self.field(nh.FIELD_OMEGA), self.Nx * self.Ny, (self.Ny, self.Nx))

    def destroy(self):
        if self.ch:
            nh.lib().nse2d_channel_destroy(self.ch)
            self.ch = None


class Les:
    """nse2d_les_t over a fine and a coarse channel descriptor."""

    def __init__(self, fine, coarse, max_pairs, max_batch, seed=1):
        L = nh.lib()
        self.fine, self.coarse = fine, coarse
        self.h = L.nse2d_les_create(ctypes.byref(fine.desc), ctypes.byref(coarse.desc), max_pairs, max_batch, seed,
                                    None)
        if not self.h:
            raise RuntimeError(nh.last_error())
        n = coarse.Nx * coarse.Ny
        self.wc, self.psic, self.tgt = nh.DeviceBuffer(n), nh.DeviceBuffer(n), nh.DeviceBuffer(n)

    def pair(self, append=True):
        """Filtered coarse state and closure target of the fine channel's current state."""
        L = nh.lib()
        rc = L.nse2d_les_pair(self.h, self.fine.field(nh.FIELD_OMEGA), self.fine.field(nh.FIELD_PSI), self.wc.ptr,
                              self.psic.ptr, self.tgt.ptr, 1 if append else 0)
        if rc:
            raise RuntimeError(nh.last_error())

    def num_pairs(self):
        return nh.lib().nse2d_les_num_pairs(self.h)

    def fit_scales(self, n=None):
        if nh.lib().nse2d_les_fit_scales(self.h, n or self.num_pairs()):
            raise RuntimeError(nh.last_error())

    def train(self, n_train, steps, batch, lr, seed=7):
        losses = (ctypes.c_float * steps)()
        r = nh.lib().nse2d_les_train(self.h, n_train, steps, batch, lr, seed, losses)
        if r < 0:
            raise RuntimeError(nh.last_error())
        return np.array(losses[:], dtype=np.float64)

    def eval(self, first, count):
        a, b = ctypes.c_double(), ctypes.c_double()
        if nh.lib().nse2d_les_eval(self.h, first, count, ctypes.byref(a), ctypes.byref(b)):
            raise RuntimeError(nh.last_error())
        return a.value, b.value

    def attach(self, ch, mode=1):
        """mode 1: closure of every stage state; 2: once per step; 0: detach."""
        if nh.lib().nse2d_les_attach(self.h, ch.ch, int(mode)):
            raise RuntimeError(nh.last_error())

    def destroy(self):
        if self.h:
            nh.lib().nse2d_les_destroy(self.h)
            self.h = None


def _rel_profile_err(U, Uref):
    return float(np.max(np.abs(U - Uref)) / np.max(np.abs(Uref)))


def _interior_rel(w, ref):
    d = w[1:-1] - ref[1:-1]
    return float(np.linalg.norm(d) / np.linalg.norm(ref[1:-1]))


class Budget:
    """K, Omega, eta and P_Omega of a state on a channel's grid (main.cu:627-629: the raw sums of
    nse2d_diag_row times 1/A, with 1/2, nu as there)."""

    def __init__(self, ch):
        self.ch = ch
        self.out = nh.DeviceBuffer(nh.DIAG_NOUT)
        self.invA = 1.0 / (ch.Nx * float(np.sum(ch.m["w_node"])))

    def __call__(self, w_ptr, psi_ptr):
        ch, c = self.ch, self.ch.c
        # <omega rhs> is not used here: omega stands in for the rhs argument
        rc = nh.lib().nse2d_diag_row(psi_ptr, w_ptr, w_ptr, ch.Nx, ch.Ny, ch.Nx / c["Lx"], (ch.Ny - 1) / 2.0, c["h"],
                                     c["F0"], c["nforce"], ch.dm["a_node"].ptr, ch.dm["a_edge"].ptr, ch.dm["y"].ptr,
                                     ch.dm["w_node"].ptr, self.out.ptr, None, None)
        if rc:
            raise RuntimeError(nh.last_error())
        o = self.out.download()
        return {"K": o[nh.DIAG_K] * self.invA, "Omega": 0.5 * o[nh.DIAG_OMEGA_W2] * self.invA,
                "eta": ch.nu * o[nh.DIAG_OMEGA_GRAD2] * self.invA, "P_Omega": o[nh.DIAG_POMEGA] * self.invA}


def _budget_errs(acc, ref):
    """Relative errors of the time-averaged budget terms."""
    return {k: abs(float(np.mean([a[k] for a in acc])) - ref[k]) / abs(ref[k]) for k in ref}


def run(fine_cfg, coarse_cfg, *, spinup_time, window_time, n_pairs, holdout, train_steps, batch, lr, n_samples,
        log=print, seed=1, compare_time=None):
    """The whole experiment; returns a dict of results.

    The DNS window [t0, t1] yields n_pairs training pairs at equal time spacing; n_samples of them
    (equally spaced, the first at t0) are also kept on the host as filtered states. The coarse runs
    start from the filtered state at t0 and land exactly on the sample times (host dt of the
    reference's rule, clipped), where they are compared with the filtered DNS: the interior
    vorticity (trajectory error), the mean profile U(y) and Re_tau. Re_tau of the filtered DNS is
    measured with the coarse grid's own wall estimator (Thom closure of the filtered state, the
    estimator an LES has), and also reported for the fine grid. Finally each coarse run is timed
    over the window with the device dt (no host round trip) for the speed comparison with the DNS.
    With compare_time > window_time the DNS goes on past the training window and the comparison
    (not the training) covers [t0, t0 + compare_time], at the same sample spacing; the statistics
    are also reported over the training window alone ("window" entries).
    """
    L = nh.lib()
    fine = Channel(fine_cfg)
    coarse_cl = Channel(coarse_cfg, ic=False)
    coarse_nc = Channel(coarse_cfg, ic=False)
    les = Les(fine, coarse_cl, n_pairs, batch, seed)
    budget = Budget(coarse_cl)
    compare_time = max(window_time, compare_time or window_time)
    out = {"fine": [fine.Nx, fine.Ny], "coarse": [coarse_cl.Nx, coarse_cl.Ny], "n_pairs": n_pairs,
           "holdout": holdout, "train_steps": train_steps, "batch": batch}
    nu_c = coarse_cl.nu
    try:
        # 1. spin-up of the DNS
        ta = time.perf_counter()
        steps_spin = 0
        while fine.time() < spinup_time:
            fine.step(200)
            steps_spin += 200
        L.nse2d_device_sync()
        out["dns_spinup_steps"] = steps_spin
        out["dns_spinup_wall_s"] = time.perf_counter() - ta
        t0 = fine.time()
        log(f"DNS spin-up to t={t0:.3f} in {steps_spin} steps, {out['dns_spinup_wall_s']:.1f} s")
        # 2. pairs, samples and DNS statistics over the window
        every = max(1, n_pairs // n_samples)
        samples = []  # (t, filtered omega, U(y) of the filtered state, Re_tau fine, Re_tau coarse estimator)
        dt_pair = window_time / n_pairs
        dns_steps, dns_wall = 0, 0.0
        k = 0
        while k < n_pairs:
            if k > 0:
                target_t = t0 + k * dt_pair
                ta = time.perf_counter()
                while fine.time() < target_t:
                    fine.step(10)
                    dns_steps += 10
                L.nse2d_device_sync()
                dns_wall += time.perf_counter() - ta
            les.pair(append=True)
            if k > 0 and k % 500 == 0:
                log(f"{k} pairs, t={fine.time():.3f}")
            if k % every == 0:
                wc = les.wc.download((coarse_cl.Ny, coarse_cl.Nx))
                re_c = math.sqrt(0.5 * nu_c * (abs(wc[0].mean()) + abs(wc[-1].mean()))) * coarse_cl.c["h"] / nu_c
                samples.append((fine.time(), wc, coarse_cl.mean_u(les.psic.ptr), fine.re_tau(), re_c,
                                budget(les.wc.ptr, les.psic.ptr)))
            k += 1
        t1 = fine.time()
        n_window = len(samples)
        # 2b. filtered samples past the training window (comparison only, no pairs kept)
        dt_sample = every * dt_pair
        while samples[-1][0] + dt_sample <= t0 + compare_time + 1e-9:
            target_t = samples[-1][0] + dt_sample
            while fine.time() < target_t:
                fine.step(10)
            les.pair(append=False)
            wc = les.wc.download((coarse_cl.Ny, coarse_cl.Nx))
            re_c = math.sqrt(0.5 * nu_c * (abs(wc[0].mean()) + abs(wc[-1].mean()))) * coarse_cl.c["h"] / nu_c
            samples.append((fine.time(), wc, coarse_cl.mean_u(les.psic.ptr), fine.re_tau(), re_c,
                            budget(les.wc.ptr, les.psic.ptr)))
        t_end = samples[-1][0]
        out["compare_window"] = [t0, t_end]
        out["dns_window"] = [t0, t1]
        out["dns_steps_window"] = dns_steps
        out["dns_wall_per_time"] = dns_wall / max(t1 - t0, 1e-30)
        dns_steps_per_time = dns_steps / max(t1 - t0, 1e-30)
        log(f"{n_pairs} pairs over t=[{t0:.3f}, {t1:.3f}], DNS {dns_steps} steps, "
            f"{out['dns_wall_per_time']:.2f} s per unit time, {len(samples)} samples")
        # 3. training
        les.fit_scales(n_pairs - holdout)
        ta = time.perf_counter()
        losses = les.train(n_pairs - holdout, train_steps, batch, lr)
        out["train_wall_s"] = time.perf_counter() - ta
        out["train_loss_first_last"] = [float(losses[:10].mean()), float(losses[-10:].mean())]
        lc, lz = les.eval(n_pairs - holdout, holdout)
        out["holdout_mse_cnn"], out["holdout_mse_zero"] = lc, lz
        log(f"trained {train_steps} steps (batch {batch}) in {out['train_wall_s']:.1f} s, loss "
            f"{losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}; hold-out normalised MSE {lc:.4f} "
            f"(zero closure {lz:.4f})")
        # 4. coarse runs from the first filtered state, compared at the sample times
        def dns_stats(sl):
            ss = samples[sl]
            return (np.mean([s[2] for s in ss], axis=0), float(np.mean([s[3] for s in ss])),
                    float(np.mean([s[4] for s in ss])),
                    {k: float(np.mean([s[5][k] for s in ss])) for k in ss[0][5]})
        U_dns, re_fine, re_filt, bud_dns = dns_stats(slice(1, None))
        U_dns_w, re_fine_w, re_filt_w, bud_dns_w = dns_stats(slice(1, n_window))
        out["re_tau_dns_fine"], out["re_tau_dns_filtered"] = re_fine, re_filt
        out["budget_dns_filtered"] = bud_dns
        res = {}
        mf, dis = nh.CLOSURE_MEAN_FREE, nh.CLOSURE_DISSIPATIVE
        for name, ch, mode, flags in (("closure", coarse_cl, 1, 0), ("closure_per_step", coarse_cl, 2, 0),
                                      ("closure_mean_free", coarse_cl, 2, mf),
                                      ("closure_dissipative", coarse_cl, 2, dis),
                                      ("closure_mean_free_dissipative", coarse_cl, 2, mf | dis),
                                      ("no_closure", coarse_nc, 0, 0)):
            if nh.lib().nse2d_les_set_constraints(les.h, flags) != 0:
                raise RuntimeError(nh.last_error())
            if mode:
                les.attach(ch, mode)
            ch.set_omega(samples[0][1], t0)
            t = t0
            errs, U_acc, re_acc, b_acc = [], [], [], []
            ok, budget_hit = True, False
            # a coarse run may take as many steps per unit time as the DNS (8x finer) before it
            # counts as destabilised
            ch.nsteps, ch.max_steps = 0, int(dns_steps_per_time * (t_end - t0)) + 1000
            t_log = t0
            try:
                for (tk, wk, _, _, _, _) in samples[1:]:
                    t = ch.advance_to(t, tk)
                    w = ch.omega()
                    if not np.all(np.isfinite(w)):
                        ok = False
                        break
                    errs.append(_interior_rel(w, wk))
                    U_acc.append(ch.mean_u())
                    re_acc.append(ch.re_tau())
                    b_acc.append(budget(ch.field(nh.FIELD_OMEGA), ch.field(nh.FIELD_PSI)))
                    if t >= t_log + 2.0:
                        t_log = t
                        log(f"{name}: t = {t:.2f}, {ch.nsteps} steps")
            except StepBudgetExceeded as ex:
                ok, budget_hit = False, True
                log(f"{name}: stopped, dt collapsed ({ex})")
            ch.max_steps = None
            nw = n_window - 1  # samples of the training window after the start
            r = {"finite": ok, "step_budget_exceeded": budget_hit, "traj_err_end": errs[-1] if errs else None,
                 "traj_err_mean": float(np.mean(errs)) if errs else None,
                 "traj_err_window_end": errs[nw - 1] if len(errs) >= nw else None}
            if ok and U_acc:
                Um, rl = np.mean(U_acc, axis=0), float(np.mean(re_acc))
                r.update(re_tau=rl, re_tau_rel_err=abs(rl - re_filt) / re_filt,
                         re_tau_rel_err_vs_fine=abs(rl - re_fine) / re_fine, U_rel_err=_rel_profile_err(Um, U_dns),
                         budget=_budget_errs(b_acc, bud_dns))
                Uw, rw = np.mean(U_acc[:nw], axis=0), float(np.mean(re_acc[:nw]))
                r["window"] = {"re_tau_rel_err": abs(rw - re_filt_w) / re_filt_w,
                               "U_rel_err": _rel_profile_err(Uw, U_dns_w),
                               "budget": _budget_errs(b_acc[:nw], bud_dns_w)}
                # Re_tau over sliding thirds of the comparison window (does the error grow?)
                n3 = max(1, len(re_acc) // 3)
                r["re_tau_rel_err_thirds"] = [
                    abs(float(np.mean(re_acc[i * n3:(i + 1) * n3 if i < 2 else None])) -
                        float(np.mean([s[4] for s in samples[1:][i * n3:(i + 1) * n3 if i < 2 else None]]))) /
                    float(np.mean([s[4] for s in samples[1:][i * n3:(i + 1) * n3 if i < 2 else None]]))
                    for i in range(3)]
            # timing over the window with the device dt (a run that destabilised is not timed)
            if budget_hit:
                r.update(timed_steps=0, wall_per_time=None)
            else:
                ch.set_omega(samples[0][1], t0)
                L.nse2d_device_sync()
                ta = time.perf_counter()
                steps = 0
                while ch.time() < t1:
                    ch.step(20)
                    steps += 20
                L.nse2d_device_sync()
                wall = time.perf_counter() - ta
                r.update(timed_steps=steps, wall_per_time=wall / max(ch.time() - t0, 1e-30))
            res[name] = r
            log(f"{name}: trajectory error {r['traj_err_end']}, Re_tau {r.get('re_tau')} vs filtered DNS "
                f"{re_filt:.2f} / fine {re_fine:.2f}, U(y) error {r.get('U_rel_err')}, budget errors "
                f"{r.get('budget')}, {r['wall_per_time']} s per unit time")
        out["runs"] = res
        def speedup(n):
            w = res[n]["wall_per_time"]
            return out["dns_wall_per_time"] / w if w else None
        out["speedup_vs_dns"] = speedup("closure")
        out["speedup_vs_dns_per_step_closure"] = speedup("closure_per_step")
        return out
    finally:
        les.destroy()
        for ch in (fine, coarse_cl, coarse_nc):
            ch.destroy()
