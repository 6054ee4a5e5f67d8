#!/usr/bin/env python3
"""Benchmark: SU(3) pure-gauge Wilson HMC at 32^4, beta = 6.0, on one MI355X.

Workload (one "step" = one HMC trajectory over the whole lattice):
  SU(3) 32^4 periodic, beta = 6.0, Haar-random start (Philox seed 2),
  50 thermalisation trajectories without Metropolis, then W warm-up and K
  timed trajectories with Metropolis; Omelyan 2nd-order integrator (2mn),
  epsilon = 0.5 (tau = 1), 12 substeps (49 steps: 25 link, 24 force).
Metric: site-updates/s = V * K / wall time of the K timed trajectories.
The headline value is the FP64 path; the FP32-mixed path (FP32 links and
momenta, FP64 sums and Metropolis) runs the same workload and is reported
beside it.  All inputs are resident in HBM when the timed region starts.

Roofline: the dominant kernel is the fused molecular-dynamics step
"force_link" (staples + TA + momentum step, then exp(t P) U into the second
field).  Its algorithmic traffic is 216 reals per site (read U and P, write P
and U', with the momenta in the trajectory's compact antihermitian layout of 9
reals per link: 1728 B FP64, 864 B FP32); its average duration comes from HIP events
recorded around every launch on the library's stream during the timed region.
Per-site algorithmic bytes of every kernel are in kernel_bytes(); the
trajectory's total is the sum over the launches actually made.

The acceptance, plaquette and <exp(-dH)> of the timed window ("timed_window")
belong to a chain only --therm + W trajectories from a Haar start, still far
from equilibrium; the "physics" objects come from the same chain continued
untimed (--equil trajectories without Metropolis, then --measure recorded
with it), near equilibrium, and are the ones to read as physics.

cpu_baseline: the CPU oracle (FP64, OpenMP) running one trajectory of the
same workload on the host cores, with the one-thread and reference-structured
(cshift copies, separate passes) figures beside it.
"""
import argparse
from ctypes import c_double as C_double
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# algorithmic bytes per site of each kernel as a trajectory runs it, R = bytes
# per real; a link is 18 reals, a momentum 18 in the matrix layout (separate
# steps, or --matrix-momenta) and 9 in the compact antihermitian one (fused
# path, the default):
#                                                compact P   matrix P
#   force_update: read U, P; write P                  -       216 R
#   force_link:   read U, P; write P, U'            216 R     288 R
#   link_update:  read U, P; write U'                 -       216 R
#   delta_h:      read U, U', P, K(8 B)           180 R + 8  216 R + 8
#   momenta:      write P, K                          -       72 R + 8
#   momenta_link: write P, K; read U; write U'    180 R + 8  216 R + 8
def kernel_bytes(name, R, compact=True):
    if compact and name in ("force_link", "delta_h", "momenta_link"):
        return {"force_link": 216 * R, "delta_h": 180 * R + 8, "momenta_link": 180 * R + 8}[name]
    return {"force_update": 216 * R, "force_link": 288 * R, "link_update": 216 * R,
            "delta_h": 216 * R + 8, "momenta": 72 * R + 8, "momenta_link": 216 * R + 8}[name]


KERNELS = ("force_link", "force_update", "link_update", "delta_h", "momenta", "momenta_link")


def trajectory_bytes_per_site(kern, R, steps, compact=True):
    """algorithmic bytes per site of one trajectory, from the launches made"""
    return sum(kernel_bytes(k, R, compact) * kern[k]["launches"] for k in KERNELS) / steps


def dominant(kern):
    return max(KERNELS, key=lambda k: (kern[k]["avg_ms"] or 0.0) * kern[k]["launches"])


def measured_traffic(prec_name, L):
    """HBM bytes per launch of the fused MD step from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/force_link_L<L>_traffic.md), or None"""
    path = os.path.join(ROOT, "profiles", f"force_link_L{L}_traffic.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == prec_name:
                return int(f[2])
    except OSError:
        pass
    return None


def run_gpu(prec, L, steps, warmup, therm, seed, scheme, eps, substeps, equil=0, measure=0):
    import numpy as np

    import mesh_amd
    from mesh_amd import gauge

    ctx = mesh_amd.default_context()
    g = mesh_amd.Grid((L, L, L, L))
    deltas = gauge.makeDeltas(scheme, eps, substeps)
    hmc = gauge.Hmc(g, prec, seed=seed, ctx=ctx)
    hmc.randomizeGaugeField()
    for _ in range(therm):
        hmc.runHmcUpdate(6.0, deltas, metropolis=False)
    for _ in range(warmup):
        hmc.runHmcUpdate(6.0, deltas)
    hmc.reset_observables()
    ctx.profile(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        hmc.runHmcUpdate(6.0, deltas)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.profile(False)
    kern = {}
    for k in KERNELS:
        ms, n = ctx.profile_read(k)
        kern[k] = dict(avg_ms=ms / n if n else None, launches=n)
    s = hmc.summary()
    out = dict(hmc=hmc, seconds=dt, V=g.size(), kernels=kern, acceptance=s["acceptance"],
               plaquette=s["plaquette"], exp_mdH=s["exp_mdH"], dH=list(hmc.deltaH_history))
    if measure > 0:
        out["physics"] = physics(hmc, deltas, therm + warmup + steps, equil, measure)
    return out


def physics(hmc, deltas, done, equil, measure):
    """The timed chain continued, untimed: `equil` trajectories without
    Metropolis (at 32^4 this integrator accepts ~10% at equilibrium, so a
    Metropolis chain would hardly move towards it), then `measure` Metropolis
    trajectories recorded.  Acceptance (against erfc(sqrt(<dH>)/2), its
    equilibrium value for Gaussian dH), plaquette, <dH> and <exp(-dH)> (= 1 at
    equilibrium, src/gauge/hmc.h:98-99) with errors from 10 blocks, the
    plaquette of the window's two halves, and the standard error
    sqrt((exp(2 <dH>) - 1) / N) that <exp(-dH)> has at equilibrium: at <dH> ~ 4
    it is far above 1 and the Creutz average cannot be checked from this
    window."""
    import numpy as np

    for _ in range(equil):
        hmc.runHmcUpdate(6.0, deltas, metropolis=False)
    hmc.reset_observables()
    for _ in range(measure):
        hmc.runHmcUpdate(6.0, deltas)

    def blocked(x):
        x = np.asarray(x, np.float64)
        b = x[: len(x) // 10 * 10].reshape(10, -1).mean(axis=1)
        return float(x.mean()), float(b.std(ddof=1) / math.sqrt(10))

    dH = np.array(hmc.deltaH_history)
    p = np.array(hmc.plaq_history)
    acc, acc_e = blocked(hmc.accept_history)
    pl, pl_e = blocked(p)
    md, md_e = blocked(dH)
    ex, ex_e = blocked(np.exp(-dH))
    h = len(p) // 2
    return {"trajectories_before": done + equil, "measured": measure,
            "acceptance": acc, "acceptance_err": acc_e,
            "acceptance_erfc": math.erfc(math.sqrt(max(md, 0.0)) / 2),
            "plaquette": pl, "plaquette_err": pl_e,
            "plaquette_first_half": float(p[:h].mean()), "plaquette_second_half": float(p[h:].mean()),
            "mean_dH": md, "mean_dH_err": md_e, "exp_minus_dH": ex, "exp_minus_dH_err_blocks": ex_e,
            "exp_minus_dH_err_equilibrium": math.sqrt(math.expm1(min(2 * max(md, 0.0), 700.0)) / len(dH))}


def same_config_compare(U64, seed, scheme, eps, substeps):
    """One trajectory from the same FP64 config and momenta in both precisions:
    plaquette and Delta H errors of FP32-mixed, with the error of Delta H split
    into that of H at the start (the FP32-rounded field and momenta) and at the
    end of the trajectory."""
    import mesh_amd
    from mesh_amd import F32, F64, gauge

    deltas = gauge.makeDeltas(scheme, eps, substeps)
    out = {}
    for key, prec, fs in ((F64, F64, 0), (F32, F32, 0), ("fsum", F32, 1)):
        h = gauge.Hmc(U64.grid, prec, seed=seed + 1000, ctx=U64.ctx)
        h.U = U64.convert(prec)
        U64.ctx.set_option("float_sums", fs)
        try:
            r = h.runHmcUpdate(6.0, deltas, metropolis=False)
        finally:
            U64.ctx.set_option("float_sums", 0)
        out[key] = (r["dH"], gauge.plaquette(h.U), r["H_old"], r["H_new"])
    return dict(dplaq=abs(out[F32][1] - out[F64][1]), ddH=abs(out[F32][0] - out[F64][0]),
                dH_f64=out[F64][0], plaq_f64=out[F64][1],
                H_old_f32_minus_f64=out[F32][2] - out[F64][2], H_new_f32_minus_f64=out[F32][3] - out[F64][3],
                # the reference's float build (--precision 1): H_old and H_new each
                # from FP32 lattice sums, then differenced
                reference_float_sums_ddH=abs(out["fsum"][0] - out[F64][0]))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(L, seed, scheme, eps, substeps):
    """The CPU oracle (FP64) on the host cores of this box.  `value`: one full
    trajectory of the workload, fused per-site restatement, every OpenMP
    thread.  Beside it, extrapolated from one pass of each stage on the same
    field (t_momenta + N_F t_force + N_L t_link + t_end): the fused restatement
    on one thread, and the reference's pass structure (cshift copies and
    separate full-lattice passes, oracle_refpass.cpp) on all threads and on one."""
    from oracle import oracle as O

    dims = [L] * 4
    V = L ** 4
    U = O.random_gauge(dims, seed)
    d = O.make_deltas(scheme, eps, substeps)
    n_force, n_link = (len(d) - 1) // 2, (len(d) + 1) // 2
    threads = O.max_threads()
    t0 = time.perf_counter()
    O.hmc_trajectory(U.copy(), dims, 6.0, d, seed, 0, metropolis=False)
    full = time.perf_counter() - t0

    def per_traj(t, refpass):
        return t[0] + n_force * t[1] + (n_link if refpass else n_link - 1) * t[2] + t[3]

    modes = {}
    for name, refpass, nt in (("fused_1_thread", False, 1), ("reference_structured", True, threads),
                              ("reference_structured_1_thread", True, 1)):
        O.set_threads(nt)
        t = O.stage_times(U, dims, 6.0, d[1], refpass=refpass)
        modes[name] = {"value": V / per_traj(t, refpass), "threads": nt,
                       "stage_seconds": [round(float(x), 4) for x in t]}
    O.set_threads(threads)
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return dict(value=V / full, unit="site-updates/s", cores=threads, kind="port",
                sample=f"one FP64 trajectory of the same workload ({L}^4, {scheme}, eps={eps}, "
                       f"{substeps} substeps) from the Haar start, fused CPU oracle, {threads} OpenMP threads; "
                       f"{full:.2f} s",
                cpu_model=cpu_model(), affinity_cpus=aff,
                modes=modes,
                modes_note=("per trajectory from one timed pass of each stage at the same size: momenta, "
                            f"{n_force} momentum steps, {n_link} link steps, end (reunitarisation, Delta H)"))


def timed_window(r, therm, warmup):
    """Acceptance, plaquette and <exp(-dH)> of the timed trajectories, labelled
    as what they are: a chain therm + warmup trajectories from a Haar start."""
    return {"acceptance": r["acceptance"], "plaquette": r["plaquette"], "exp_minus_dH": r["exp_mdH"],
            "note": f"timed trajectories {therm + warmup}.. of a chain from a Haar start, before "
                    "equilibrium (exp_minus_dH far from 1); the 'physics' object holds the equilibrium values"}


def kernel_only(name, prec, L, reps, variant=-1):
    """Profiling mode: `reps` launches of one kernel on a seeded L^4 field,
    preceded by `reps` plain copies of one gauge field (a transfer of known
    size, used to calibrate rocprof byte counters).  Prints one JSON line."""
    import mesh_amd
    from mesh_amd import gauge

    ctx = mesh_amd.default_context()
    ctx.set_option("force_variant", variant)
    g = mesh_amd.Grid((L,) * 4)
    U = mesh_amd.GaugeField(g, prec, ctx)
    U2 = mesh_amd.GaugeField(g, prec, ctx)
    P = mesh_amd.GaugeField(g, prec, ctx)
    K = mesh_amd.DeviceArray(g.size(), ctx)
    gauge.randomGaugeField(U, 2)
    gauge.randomMomenta(P, 2, 0, K)
    for _ in range(reps):
        ctx.call("mesh_stream_copy", U2.ptr, U.ptr, U.bytes)
    ctx.synchronize()
    ctx.profile(True)
    if name == "force_link":
        # the fused step only runs inside trajectories: time `reps` of them
        h = gauge.Hmc(g, prec, seed=2, ctx=ctx)
        h.U = U
        d = gauge.makeDeltas("2mn", 0.5, 12)
        for _ in range(reps):
            h.runHmcUpdate(6.0, d, metropolis=False)
    for _ in range(reps if name != "force_link" else 0):
        if name == "force_update":
            gauge.momentumStep(U, P, 6.0, 1e-6)
        elif name == "link_update":
            gauge.linkStep(U2, U, P, 1e-6)
        elif name == "delta_h":
            gauge.deltaH(U, U2, P, K, 6.0)
        elif name == "landau_or":
            ctx.call("mesh_landau_or_halfsweep", g.ref, U.prec, U.ptr, None, 0, C_double(1.7))
        else:
            raise SystemExit(f"unknown kernel {name}")
    ctx.synchronize()
    ms, n = ctx.profile_read(name)
    print(json.dumps({"kernel": name, "prec": "f64" if prec == 0 else "f32", "L": L, "variant": variant,
                      "avg_ms": ms / n,
                      "launches": n, "copy_bytes_per_launch": 2 * U.bytes}), flush=True)


def u1_ess(L, beta, batch, train_steps, hmc_steps, flow_steps, seed, flows=("pflow", "pcflow", "coupling"),
           substeps=20, start="hot", n_therm=500, layers=8):
    """U(1) 2-d at L x L: ESS/s of the topological charge Q for plain HMC
    (2lf, tau = 1, `substeps` steps) and for HMC interleaved 1:1 with
    independence MH steps proposed by each trained flow ("pflow": plaquette
    flow, "coupling": `layers` coupling layers on links, "pcflow": `layers`
    coupling layers on plaquettes).  start: "hot" (uniform angles) or
    "cold" (unit field) for every chain.  Training time is reported apart
    from the sampling time the ESS/s figures use."""
    import numpy as np

    from mesh_amd import gauge
    from mesh_amd import u1 as U1

    deltas = gauge.makeDeltas("2lf", 0.5, substeps)
    out = {"metric": f"U(1) 2-d ESS/s of Q at L={L} beta={beta}, flow MH + HMC vs HMC",
           "config": {"workload": f"U(1) {L}x{L} beta={beta} batch={batch}, HMC 2lf tau=1 {substeps} steps, "
                                  f"{start} start, flow MH 1:1", "L": L, "beta": beta, "batch": batch,
                      "hmc_substeps": substeps, "start": start},
           "dtype": "f32 angles, f64 energies and log-weights", "data": "synthetic: Philox-seeded starts"}

    def summary(r, extra):
        ess, tau, err = U1.ess_per_second(r)
        taup, _ = U1.tau_int(r["plaq"])
        frozen = not math.isfinite(tau)
        d = {"ess_per_s": ess, "tau_int_Q": None if frozen else tau, "Q_frozen": frozen, "tau_err": err,
             "tau_int_plaq": taup,
             "acceptance_hmc": r["acc_hmc"], "seconds": r["seconds"], "steps": r["Q"].shape[1],
             "chi_t": float(np.mean(r["Q"] ** 2) / (L * L)), "mean_cos": float(r["plaq"].mean()),
             "tau_window_capped": bool(frozen or 2 * tau >= r["Q"].shape[1] / 2)}
        d.update(extra)
        return d

    def new_ensemble(s):
        e = U1.U1Ensemble(L, L, batch)
        if start == "cold":
            e.cold_start()
        else:
            e.random_start(s)
        return e

    ens = new_ensemble(seed)
    out["hmc"] = summary(U1.run_chains(ens, beta, deltas, hmc_steps, seed, n_therm=n_therm), {})
    for kind in flows:
        if kind == "pflow":
            flow = U1.PlaqFlow(L, L)
            lr, lr_final = 2e-2, 1e-3
        elif kind == "pcflow":
            flow = U1.PlaqCouplingFlow(L, L, seed=seed, layers=layers)
            lr, lr_final = 1e-3, 1e-4
        else:
            flow = U1.Flow(L, L, seed=seed, layers=layers)
            lr, lr_final = 1e-3, 1e-4
        t0 = time.perf_counter()
        hist = flow.train(beta, train_steps, batch=batch, lr=lr, lr_final=lr_final, seed=seed + 1)
        flow.ctx.synchronize()
        t_train = time.perf_counter() - t0
        ens2 = new_ensemble(seed + 2)
        r2 = U1.run_chains(ens2, beta, deltas, flow_steps, seed + 3, flow=flow, n_therm=100)
        out[kind] = summary(r2, {"acceptance_flow": r2["acc_flow"], "train_steps": train_steps,
                                 **({"layers": layers} if kind in ("coupling", "pcflow") else {}),
                                 "train_seconds": t_train, "loss_first": float(np.mean(hist[:10])),
                                 "loss_last": float(np.mean(hist[-50:]))})
        out[kind]["ratio_vs_hmc"] = out[kind]["ess_per_s"] / out["hmc"]["ess_per_s"]
    out["value"] = max(out[k]["ratio_vs_hmc"] for k in flows)
    out["unit"] = "ratio of ESS/s of Q (best flow / HMC)"
    out["higher_is_better"] = True
    return out


# SU(2) algorithmic bytes per site (R bytes per real, 4 directions of 4-real
# quaternions; momenta stored as quaternions with w = 0):
#   force_update: read U, P; write P     48 R
#   link_update:  read U, P; write U'    48 R
#   momenta:      write P, K             16 R + 8
#   delta_h:      read U, U', P, K       48 R + 8
#   landau_or:    one half-sweep reads and writes every link once (each link
#                 joins a site of each parity) and the transformation of
#                 half the sites: 32 R + 4 R
#   force_link:   the two fused (the trajectory's default): read U, P;
#                 write P, U'                                       64 R
SU2_BYTES = {"force_update": (48, 0), "link_update": (48, 0), "momenta": (16, 8), "delta_h": (48, 8),
             "landau_or": (36, 0), "force_link": (64, 0)}


def su2_bench(L, steps, warmup, therm, seed, beta=2.3, scheme="2mn", eps=0.5, substeps=12, landau_sweeps=2000):
    """SU(2) pure-gauge HMC (Hmc<SU2<T>>) at L^4 in FP64 and FP32: trajectories
    per second as site-updates/s, each kernel's average launch time (HIP events
    on the library's stream) and its fraction of 8 TB/s on the algorithmic
    bytes above; then Landau overrelaxation of the thermalised field to
    theta < 1e-12 (or landau_sweeps sweeps)."""
    import mesh_amd
    from mesh_amd import F32, F64, gauge
    from mesh_amd import su2 as S2

    ctx = mesh_amd.default_context()
    g = mesh_amd.Grid((L,) * 4)
    V = g.size()
    d = gauge.makeDeltas(scheme, eps, substeps)
    out = {"metric": f"SU(2) HMC site-updates/s at {L}^4", "unit": "site-updates/s", "higher_is_better": True,
           "config": {"workload": f"SU(2) {L}^4 beta={beta} {scheme} eps={eps} {substeps} substeps, "
                                  f"{therm} thermalisation trajectories", "L": L, "beta": beta},
           "data": "synthetic: Philox-seeded Haar start"}
    for prec, name in ((F64, "f64"), (F32, "f32")):
        R = 8 if prec == F64 else 4
        h = S2.Hmc(g, prec, seed=seed, ctx=ctx)
        h.randomizeGaugeField()
        for _ in range(therm):
            h.runHmcUpdate(beta, d, metropolis=False)
        for _ in range(warmup):
            h.runHmcUpdate(beta, d)
        h.accept_history.clear()
        h.plaq_history.clear()
        ctx.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(steps):
            h.runHmcUpdate(beta, d)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        kern = {}
        for k in ("force_link", "force_update", "link_update", "momenta", "delta_h"):
            ms, n = ctx.profile_read(k)
            a, b = SU2_BYTES[k]
            avg = ms / n if n else None
            kern[k] = {"avg_ms": avg, "launches": n,
                       "frac_of_hbm": (a * R + b) * V / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS if avg else None}
        ctx.profile(False)
        lan = S2.Landau(h.U)
        ctx.profile(True)
        ctx.profile_reset()
        t1 = time.perf_counter()
        sweeps, theta = lan.run(landau_sweeps, omega=1.7, tol=1e-12 if prec == F64 else 1e-9)
        ctx.synchronize()
        tl = time.perf_counter() - t1
        ms, n = ctx.profile_read("landau_or")
        ctx.profile(False)
        half = ms / n if n else None
        out[name] = {"value": V * steps / dt, "ms_per_step": 1e3 * dt / steps,
                     "acceptance": float(sum(h.accept_history) / len(h.accept_history)),
                     "plaquette": float(sum(h.plaq_history) / len(h.plaq_history)), "kernels": kern,
                     "landau": {"sweeps": sweeps, "theta": theta, "seconds": tl,
                                "site_sweeps_per_s": V * sweeps / tl, "half_sweep_ms": half,
                                "frac_of_hbm": (SU2_BYTES["landau_or"][0] * R * V) / (half * 1e-3) / 1e9
                                / HBM_PEAK_GBS if half else None}}
    out["value"] = out["f64"]["value"]
    out["dtype"] = "f64 (f32 beside it)"
    return out


def u1_landau_bench(L, seed, therm, max_sweeps, beta=1.0, omega=1.8, tol=1e-12):
    """Landau gauge fixing of a U(1) L^4 field (Landau<U1<T>>, FP32 angles)
    thermalised by `therm` HMC trajectories: overrelaxation sweeps to
    theta < tol, the half-sweep time from HIP events and its fraction of 8 TB/s
    on 36 B per site (every link read and written once per half-sweep, the
    phases of half the sites)."""
    import mesh_amd
    from mesh_amd import gauge
    from mesh_amd import u1 as U1

    ctx = mesh_amd.default_context()
    dims = (L,) * 4
    ens = U1.U1Ensemble(None, None, 1, ctx, dims=dims)
    ens.random_start(seed)
    d = gauge.makeDeltas("2lf", 0.5, 20)
    for t in range(therm):
        ens.hmc(beta, d, seed, t, metropolis=False)
    g = mesh_amd.Grid(dims)
    V = g.size()
    U = U1.U1Field(g, None, ctx).upload(ens.angles()[0])
    lan = U1.Landau(U)
    F0 = lan()
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    sweeps, theta = lan.run(max_sweeps, omega=omega, tol=tol)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms, n = ctx.profile_read("landau_or")
    ctx.profile(False)
    half = ms / n if n else None
    return {"metric": f"U(1) Landau gauge fixing site-sweeps/s at {L}^4", "unit": "site-sweeps/s",
            "value": V * sweeps / dt, "higher_is_better": True,
            "config": {"workload": f"U(1) {L}^4 beta={beta} after {therm} HMC trajectories; overrelaxation "
                                   f"omega={omega} to theta < {tol:g}", "L": L},
            "sweeps": sweeps, "theta": theta, "functional_start": F0, "functional_end": lan(),
            "seconds": dt, "half_sweep_ms": half,
            "frac_of_hbm": 36 * V / (half * 1e-3) / 1e9 / HBM_PEAK_GBS if half else None,
            "dtype": "f32 angles, f64 sums", "data": "synthetic: Philox-seeded start, HMC-thermalised"}


def landau_bench(L, seed, therm, max_sweeps, omega=1.7, tol=1e-12, ascent_iters=0):
    """Landau gauge fixing of a thermalised SU(3) L^4 configuration (HMC as in
    the main workload, `therm` trajectories) by even/odd overrelaxation, in
    FP32 links (the measured path) and FP64; theta checked every 10 sweeps.
    With ascent_iters > 0 also the reference's line-search solver (FP64, at
    most that many iterations) and FP64 overrelaxation, both to the
    reference's stopping condition (condition() < 1e-10, theta < 4e-10)."""
    import mesh_amd
    from mesh_amd import F32, F64, gauge

    ctx = mesh_amd.default_context()
    g = mesh_amd.Grid((L,) * 4)
    V = g.size()
    hmc = gauge.Hmc(g, F64, seed=seed, ctx=ctx)
    hmc.randomizeGaugeField()
    d = gauge.makeDeltas("2mn", 0.5, 12)
    for _ in range(therm):
        hmc.runHmcUpdate(6.0, d, metropolis=False)
    out = {"metric": f"Landau gauge fixing site-sweeps/s at {L}^4 (FP32 links)",
           "config": {"workload": f"SU(3) {L}^4 beta=6.0 after {therm} HMC trajectories; overrelaxation "
                                  f"omega={omega}, theta < {tol:g} checked every 10 sweeps", "L": L},
           "data": "synthetic: Philox-seeded Haar start, thermalised by HMC"}
    for prec in (F32, F64):
        R = 4 if prec == F32 else 8
        U = hmc.U.convert(prec) if prec == F32 else hmc.U.clone()
        lan = gauge.Landau(U)
        ctx.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        sweeps, theta = lan.run(max_sweeps, omega=omega, tol=tol, check_every=10)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ctx.profile(False)
        ms, n = ctx.profile_read("landau_or")
        # per half-sweep: every link read and written once, the transformation of
        # the updated half read and written: (144 + 18) R per site
        alg = 162 * R * V
        name = "f32" if prec == F32 else "f64"
        out[name] = {"sweeps": sweeps, "theta": theta, "functional": lan(), "plaquette": gauge.plaquette(U),
                     "seconds": dt, "site_sweeps_per_s": V * sweeps / dt,
                     "half_sweep_ms": ms / max(n, 1),
                     "half_sweep_frac_of_hbm": alg / (ms / max(n, 1) * 1e-3) / (HBM_PEAK_GBS * 1e9)}
    if ascent_iters > 0:
        cmp = {}
        U = hmc.U.clone()
        lan = gauge.Landau(U)
        ctx.synchronize()
        t0 = time.perf_counter()
        it, cond = lan.run_ascent(ascent_iters, tol=1e-10)
        ctx.synchronize()
        cmp["reference_ascent"] = {"iterations": it, "condition": cond, "converged": bool(cond < 1e-10),
                                   "functional": lan(), "seconds": time.perf_counter() - t0}
        U = hmc.U.clone()
        lan = gauge.Landau(U)
        ctx.synchronize()
        t0 = time.perf_counter()
        sweeps, theta = lan.run(max_sweeps, omega=omega, tol=4e-10, check_every=10)
        ctx.synchronize()
        cmp["overrelaxation"] = {"sweeps": sweeps, "condition": theta / 4, "functional": lan(),
                                 "seconds": time.perf_counter() - t0}
        out["f64_to_reference_condition"] = cmp
    out["plaquette_before"] = gauge.plaquette(hmc.U)
    out["value"] = out["f32"]["site_sweeps_per_s"]
    out["unit"] = "site-sweeps/s"
    out["higher_is_better"] = True
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--kernel-only", default=None,
                    help="profiling mode: run one kernel (force_link, force_update, link_update, delta_h, "
                         "landau_or)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--force-variant", type=int, default=-1,
                    help="kernel-only mode: force kernel variant (-1 library default, 0, 1, 2)")
    ap.add_argument("--prec", default="f64")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--therm", type=int, default=50)
    ap.add_argument("--equil", type=int, default=300,
                    help="untimed trajectories without Metropolis after the timed ones, before the physics window")
    ap.add_argument("--measure", type=int, default=100, help="trajectories of the physics window (0: none)")
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--scheme", default="2mn")
    ap.add_argument("--epsilon", type=float, default=0.5)
    ap.add_argument("--substeps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32-mixed run")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="library tuning option (mesh_set_option) for the run, e.g. xcd_remap=4; repeatable")
    ap.add_argument("--matrix-momenta", action="store_true",
                    help="keep the trajectory's momenta as 18-real matrices (option compact_momenta = 0; "
                         "same bits, more traffic)")
    ap.add_argument("--landau", action="store_true",
                    help="instead: Landau gauge fixing of a thermalised 32^4 configuration (one JSON line)")
    ap.add_argument("--landau-sweeps", type=int, default=30000)
    ap.add_argument("--landau-ascent", type=int, default=0, metavar="N",
                    help="with --landau: also the reference's line-search solver (at most N iterations) "
                         "against overrelaxation, both to condition() < 1e-10")
    ap.add_argument("--su2", action="store_true", help="SU(2) HMC and Landau at --L (its own JSON line)")
    ap.add_argument("--u1-landau", action="store_true", help="U(1) Landau gauge fixing at --L^4 (its own JSON line)")
    ap.add_argument("--u1", action="store_true",
                    help="instead: U(1) 2-d ESS/s of Q, flow-assisted vs plain HMC (one JSON line)")
    ap.add_argument("--u1-L", type=int, default=32)
    ap.add_argument("--u1-beta", type=float, default=6.0)
    ap.add_argument("--u1-batch", type=int, default=256)
    ap.add_argument("--u1-train", type=int, default=4000)
    ap.add_argument("--u1-hmc-steps", type=int, default=4000)
    ap.add_argument("--u1-flow-steps", type=int, default=1000)
    ap.add_argument("--u1-flows", default="pflow,pcflow,coupling", help="flows to measure (pflow, pcflow, coupling)")
    ap.add_argument("--u1-substeps", type=int, default=20, help="HMC steps per unit trajectory")
    ap.add_argument("--u1-start", choices=("hot", "cold"), default="hot")
    ap.add_argument("--u1-therm", type=int, default=500, help="HMC thermalisation trajectories")
    ap.add_argument("--u1-layers", type=int, default=8, help="coupling layers of the coupling flow")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {a.gpus} is not supported "
              "(the lattice is not domain-decomposed)", file=sys.stderr)
        sys.exit(2)

    import mesh_amd
    from mesh_amd import F32, F64

    options = {}
    for o in a.option:
        name, _, v = o.partition("=")
        options[name] = int(v)
        mesh_amd.default_context().set_option(name, int(v))
    if a.landau:
        print(json.dumps(landau_bench(a.L, a.seed, a.therm, a.landau_sweeps, ascent_iters=a.landau_ascent)),
              flush=True)
        return
    if a.u1_landau:
        print(json.dumps(u1_landau_bench(a.L, a.seed, a.therm, 20000)), flush=True)
        return
    if a.su2:
        print(json.dumps(su2_bench(a.L, a.steps, a.warmup, a.therm, a.seed)), flush=True)
        return
    if a.u1:
        print(json.dumps(u1_ess(a.u1_L, a.u1_beta, a.u1_batch, a.u1_train, a.u1_hmc_steps, a.u1_flow_steps,
                                a.seed, tuple(a.u1_flows.split(",")), a.u1_substeps, a.u1_start,
                                a.u1_therm, a.u1_layers)), flush=True)
        return
    if a.kernel_only:
        kernel_only(a.kernel_only, F64 if a.prec == "f64" else F32, a.L, a.reps, a.force_variant)
        return

    compact = not a.matrix_momenta
    mesh_amd.default_context().set_option("compact_momenta", int(compact))
    momenta = ("compact antihermitian, 9 reals per link (bit-identical to 18-real matrices)" if compact
               else "18-real matrices")
    r64 = run_gpu(F64, a.L, a.steps, a.warmup, a.therm, a.seed, a.scheme, a.epsilon, a.substeps, a.equil,
                  a.measure)
    V = r64["V"]
    value = V * a.steps / r64["seconds"]
    dom = dominant(r64["kernels"])
    fk = r64["kernels"][dom]
    R = 8
    alg = kernel_bytes(dom, R, compact) * V
    achieved = alg / (fk["avg_ms"] * 1e-3) / 1e9
    line = {
        "metric": "SU(3) HMC site-updates/s at 32^4 (FP64)",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * r64["seconds"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Philox-seeded Haar-random SU(3) start, seed 2, thermalised",
        "config": {"workload": f"SU(3) Wilson HMC {a.L}^4 beta=6.0 {a.scheme} eps={a.epsilon} "
                               f"substeps={a.substeps} (tau=1), FP64",
                   "L": a.L, "beta": 6.0, "thermalisation": a.therm, "momenta": momenta,
                   **({"options": options} if options else {})},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("f64", a.L) if dom == "force_link" and compact else None,
                     "kernel": dom, "avg_ms": fk["avg_ms"],
                     "algorithmic_bytes_per_launch": alg},
        "trajectory_roofline_frac": value * trajectory_bytes_per_site(r64["kernels"], 8, a.steps, compact) / 8e12,
        "timed_window": timed_window(r64, a.therm, a.warmup),
        **({"physics": r64["physics"]} if "physics" in r64 else {}),
        "kernels_f64": r64["kernels"],
    }
    if not a.no_f32:
        r32 = run_gpu(F32, a.L, a.steps, a.warmup, a.therm, a.seed, a.scheme, a.epsilon, a.substeps, a.equil,
                      a.measure)
        v32 = V * a.steps / r32["seconds"]
        dom32 = dominant(r32["kernels"])
        fk32 = r32["kernels"][dom32]
        ach32 = kernel_bytes(dom32, 4, compact) * V / (fk32["avg_ms"] * 1e-3) / 1e9
        cmp = same_config_compare(r64["hmc"].U, a.seed, a.scheme, a.epsilon, a.substeps)
        line["f32_mixed"] = {
            "value": v32, "ms_per_step": 1e3 * r32["seconds"] / a.steps, "ratio_vs_f64": v32 / value,
            "timed_window": timed_window(r32, a.therm, a.warmup),
            **({"physics": r32["physics"]} if "physics" in r32 else {}),
            "dominant_kernel": dom32, "dominant_frac_of_hbm": ach32 / HBM_PEAK_GBS,
            "dominant_traffic": measured_traffic("f32", a.L) if dom32 == "force_link" and compact else None,
            "trajectory_roofline_frac": v32 * trajectory_bytes_per_site(r32["kernels"], 4, a.steps, compact) / 8e12,
            "same_config_vs_f64": cmp, "kernels": r32["kernels"],
        }
    if not a.no_cpu:
        cb = cpu_baseline(a.L, a.seed, a.scheme, a.epsilon, a.substeps)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
