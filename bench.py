#!/usr/bin/env python3
"""Throughput of the RPMD hot path on one MI355X.

Workload (the batched recrossing-children configuration): CH4 + H on the CH5
surface, 4096 trajectories x 32 beads, centroid constrained to the dividing
surface xi = 1, Maxwell-Boltzmann momenta at 500 K in +-p pairs, FP64, dt =
0.1 fs, child steps with the reaction coordinate of the centroid and the
recrossing-factor numerator every step, all steps of a run in one HIP graph.
A step is one velocity-Verlet RPMD step of every trajectory.

Printed: one JSON line.  value = bead-force-evals/s of the batched FP64 run
(inputs resident in HBM, schedule chosen by the engine's autotuner).  Also
reported: the same batch in mixed precision (FP32 force arithmetic, FP64
state and accumulation) with its force error and energy drift, the plain
FP64 one-trajectory-per-launch form, the roofline of the step kernel, a CPU
baseline, and the other workloads: a ragged batch of 8192 trajectories (H3
and CH4H, 4..64 beads), a 3000-atom box with 16 beads (Andersen NVT) and 50
Adam steps of force matching on that box with 64 beads.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 50] [--skip-extras]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work (oracle/flopcount.cpp, the reference's own arithmetic):
# +,-,*,/ per bead-force-evaluation of the CH5 surface; elementary functions
# are counted separately and not included in the flop figure.
CH4H_FLOPS_PER_BEAD = 12707
CH4H_TRANSCENDENTALS_PER_BEAD = 321
# the box force field (oracle/ff_box.c through the same counter): one
# intermolecular pair inside the cutoff (energy and gradient, the minimum
# image included), one bond, one angle
BOX_FLOPS = {"pair_in": 95, "bond": 49, "angle": 116}
FP64_PEAK_TFLOPS = 78.6            # MI355X dense FP64 (vector and matrix)
# HBM bytes of one launch of the folded child step kernel at 4096
# trajectories x 32 beads, from rocprofv3 FETCH_SIZE and WRITE_SIZE counters
# scaled by a plain 1 GiB copy (profiles/fold_step_kernel_ch4h_children_fp64.md):
# a fixed part (q, p and the gradient read and written once per launch) and
# a part per step (the child centroids of every step, 3 x 6 doubles per
# trajectory), from launches of 50 and 100 steps (145.6 and 176.2 MB).
MEASURED_TRAFFIC = {("folded", 4096, 32): (115.0e6, 0.612e6)}


def launch_traffic(nmk_name, ntraj, nbeads, spl):
    t = MEASURED_TRAFFIC.get((nmk_name, ntraj, nbeads))
    return None if t is None else t[0] + spl * t[1]


NMK_NAMES = {1: "valu", 2: "mfma", 3: "folded"}


def step_flops(natoms, nbeads, folded=False):
    """Algorithmic flops of one RPMD step of one trajectory: two half kicks
    (2 flops per coordinate each), the normal-mode transform pair on p and q
    (24 N n^2) plus the per-mode propagator (6 per coordinate), and the surface.
    Folded (the reference transform keeps the beads j and n - j equal): the
    same on the n/2 + 1 distinct beads, with the transform pair on the
    (n/2) x (n/2) block plus its rank-1 row and column (2 h^2 + 4 h per column)."""
    if not folded:
        nc = 3 * natoms * nbeads
        return 4 * nc + 24 * natoms * nbeads ** 2 + 6 * nc + nbeads * CH4H_FLOPS_PER_BEAD
    h = nbeads // 2
    nc = 3 * natoms * (h + 1)
    return 4 * nc + 12 * natoms * (2 * h * h + 4 * h) + 6 * nc + (h + 1) * CH4H_FLOPS_PER_BEAD


def step_flops_parts(natoms, nbeads, folded=False):
    """The same work split into the surface, the normal-mode transform pair
    and the rest (kicks and the per-mode propagator); the three run fused in
    one kernel, so they share one time."""
    n = nbeads // 2 + 1 if folded else nbeads
    surface = n * CH4H_FLOPS_PER_BEAD
    rest = 10 * 3 * natoms * n
    return {"surface": surface, "normal_modes": step_flops(natoms, nbeads, folded) - surface - rest, "rest": rest}


def prepare_children(cam, workloads, ntraj, nbeads, precision, steps_per_launch=0, seed=0,
                     nm_kernel=None, schedule=0, waves_per_eu=0):
    w = workloads.ch4h_children(ntraj=ntraj, nbeads=nbeads, seed=seed)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nbeads, ntraj, w["dt"], w["masses"], precision=precision,
                       ensemble=cam.CAM_ENS_CHILD, reaction=w["reaction"], seed=seed + 17,
                       steps_per_launch=steps_per_launch, schedule=schedule, waves_per_eu=waves_per_eu,
                       nm_kernel=cam.CAM_NMK_AUTO if nm_kernel is None else nm_kernel)
    e.set_state(w["q"], w["p"], w["beta"])
    st = e.constrain(project_p=False)
    if st.any():
        raise RuntimeError(f"{int(st.sum())} trajectories failed the centroid constraint")
    e.draw_momenta(draw=1, antithetic_pairs=True)
    e.begin_children()
    return e, w


def timed_run(e, steps, warmup):
    e.prepare(warmup)
    e.prepare(steps)
    e.run(warmup)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.run(steps, sync=False)
    e.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_launch, spl, nl = e.last_timing()
    return wall, ms_launch, spl, nl


def force_error_mixed(cam, workloads, ntraj, nbeads):
    """Mixed vs FP64 forces on the workload's start state: max |dF| / max |F|
    over the batch, the worst per-bead |dF|_inf / |F|_inf, and the worst
    per-bead |dF|_inf / max(|F|_inf, 1e-3 hartree/bohr)."""
    g = {}
    for prec in (cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED):
        e, _ = prepare_children(cam, workloads, ntraj, nbeads, prec)
        g[prec] = e.get_forces()[1]
        e.close()
    a, b = g[cam.CAM_PREC_FP64], g[cam.CAM_PREC_MIXED]
    num = np.abs(a - b).reshape(a.shape[0] * a.shape[1], -1).max(axis=1)
    den = np.abs(a).reshape(a.shape[0] * a.shape[1], -1).max(axis=1)
    return float(num.max() / den.max()), float(np.max(num / den)), float(np.max(num / np.maximum(den, 1e-3)))


def energy_drift(cam, workloads, nsteps=20000):
    """NVE drift of the ring-polymer energy (hartree per ps per trajectory):
    64 CH4H trajectories of 32 beads, orthogonal normal modes, 2 ps at 0.1 fs,
    FP64 and mixed.  Per trajectory, the slope of a least-squares line through
    H(t) sampled every 100 steps; reported as the mean of |slope| and as the
    slope of the ensemble-mean energy."""
    out = {}
    for name, prec in (("fp64", cam.CAM_PREC_FP64), ("mixed", cam.CAM_PREC_MIXED)):
        w = workloads.ch4h_children(ntraj=64, nbeads=32, seed=5)
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 32, 64, w["dt"], w["masses"], precision=prec,
                           nm_mode=cam.CAM_NM_ORTHOGONAL)
        rng = np.random.default_rng(5)
        p = rng.normal(size=w["q"].shape) * np.sqrt(w["masses"][None, None, :, None] * 32 / w["beta"][:, None, None, None])
        e.set_state(w["q"], p, w["beta"])
        chunk = 100
        ts, hs = [], []
        for i in range(nsteps // chunk + 1):
            o = e.observables()
            ts.append(i * chunk * w["dt"])
            hs.append(o["epot"] + o["ekin"] + o["espring"])
            if i < nsteps // chunk:
                e.run(chunk)
        e.close()
        hs = np.array(hs)
        t_ps = np.array(ts) * 2.4188843265857e-5
        slopes = np.polyfit(t_ps, hs - hs[0], 1)[0]
        out[name] = {"mean_abs_drift_eh_per_ps": float(np.abs(slopes).mean()),
                     "ensemble_drift_eh_per_ps": float(np.polyfit(t_ps, (hs - hs[0]).mean(axis=1), 1)[0]),
                     "rms_fluctuation_eh": float((hs - hs[0]).std(axis=0).mean()), "t_ps": float(t_ps[-1])}
    return out


def ragged_batch(cam, workloads, steps=200):
    """8192 trajectories, half H3 and half CH4H, nbeads uniform on {4..64},
    T uniform on [200, 1000] K, NVE, as one batch: one engine per bucket, a
    run one captured graph with a branch per bucket.  Throughput in FP64 and mixed; mixed-vs-FP64 force error per
    (surface, nbeads) bucket on the start state."""
    import torch
    buckets = workloads.ragged(ntraj=8192, seed=1)
    beads = sum(len(d["beta"]) * nb for _, nb, d in buckets)
    out, forces = {}, {}
    for name, prec in (("fp64", cam.CAM_PREC_FP64), ("mixed", cam.CAM_PREC_MIXED)):
        rb = cam.RaggedBatch(buckets, precision=prec)
        rb.set_states()
        forces[name] = [e.get_forces()[1] for e in rb.engines]
        rb.autotune(20)
        for e in rb.engines:
            e.prepare(steps)
        rb.run(steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms = rb.run(steps)
        wall = time.perf_counter() - t0
        rb.close()
        out[name] = {"bead_force_evals_per_s": beads * steps / wall, "ms_gpu": ms, "wall_s": wall}
    err = {}
    for i, (surf, nb, _) in enumerate(buckets):
        a, b = forces["fp64"][i], forces["mixed"][i]
        num = np.abs(a - b).reshape(a.shape[0] * a.shape[1], -1).max(axis=1)
        den = np.abs(a).reshape(a.shape[0] * a.shape[1], -1).max(axis=1)
        err[f"{'H3' if surf == cam.CAM_SURFACE_H3 else 'CH4H'}/{nb}"] = float(num.max() / den.max())
    out["rel_force_err_per_bucket"] = err
    out["trajectories"], out["beads"], out["steps"] = 8192, beads, steps
    return out


def box_nvt(cam, steps=500):
    """3000-atom box (1000 bent triatomics, 31 angstrom, 10 angstrom cutoff),
    16 beads, Andersen NVT at 300 K (all beads redrawn every 100 steps),
    0.1 fs: bead-force-evals/s where one evaluation is one bead of the box."""
    import torch
    from caracal_amd import ff as cff, ffbox
    s = ffbox.box_system(nmol=1000, box_angstrom=31.0, seed=2)
    ff = cff.ForceField(s, rc_angstrom=10.0)
    m = ffbox.masses(s["type"])
    beta = float(cam.beta_of(300.0))
    nb = 16
    rng = np.random.default_rng(1)
    q = s["x"][None] + rng.normal(0, 0.02, (nb,) + s["x"].shape)
    p = rng.normal(size=q.shape) * np.sqrt(m[None, :, None] * nb / beta)
    out = {}
    for name, prec in (("fp64", cam.CAM_PREC_FP64), ("mixed", cam.CAM_PREC_MIXED)):
        eng = cff.BoxEngine(ff, nb, 0.1 * cam.FS_TO_AU, m, beta, ffbox.theta_ref(), precision=prec,
                            nm_mode=cam.CAM_NM_REFERENCE, andersen_step=100, seed=5)
        eng.set_state(q, p)
        eng.prepare(steps)
        eng.run(10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms = eng.run(steps)
        wall = time.perf_counter() - t0
        eng.close()
        out[name] = {"bead_force_evals_per_s": nb * steps / wall, "ms_per_step": wall / steps * 1e3, "ms_gpu": ms}
    out["atoms"], out["nbeads"], out["steps"] = len(s["type"]), nb, steps
    ff.close()
    # algorithmic work of one step: every intermolecular pair inside the
    # cutoff once (Newton's third law; counted at the start geometry),
    # bonds, angles and the normal-mode step of the box (24 N n^2).  Pairs
    # beyond the cutoff cost nothing here: the rows never hold most of them.
    # The kernel evaluates each pair from both of its atoms (fixed-order sums
    # without atomics), so the pair flops it executes are twice the counted
    # ones ("evaluated").
    x = s["x"]
    n = len(x)
    box = s["box"]
    inter = 0
    inside = 0
    rc = 10.0 / 0.52917721092
    for i in range(n - 1):
        d = x[i + 1:] - x[i]
        d -= box * np.rint(d / box)
        r2 = (d * d).sum(axis=1)
        other = s["mol"][i + 1:] != s["mol"][i]
        inter += int(other.sum())
        inside += int((other & (r2 < rc * rc)).sum())
    per_bead = (inside * BOX_FLOPS["pair_in"] + len(s["bonds"]) * BOX_FLOPS["bond"]
                + len(s["angles"]) * BOX_FLOPS["angle"])
    flops = nb * per_bead + 24 * n * nb * nb
    evaluated = flops + nb * inside * BOX_FLOPS["pair_in"]
    out["pairs_per_bead"] = {"intermolecular": inter, "inside_cutoff": inside}
    out["flops_per_step"] = flops
    out["evaluated_flops_per_step"] = evaluated
    out["roofline"] = {name: {"achieved": flops / (out[name]["ms_per_step"] * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                              "unit": "TFLOP/s",
                              "frac": flops / (out[name]["ms_per_step"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                              "frac_evaluated": evaluated / (out[name]["ms_per_step"] * 1e-3) / 1e12
                              / FP64_PEAK_TFLOPS}
                       for name in ("fp64", "mixed")}
    out["roofline"]["note"] = ("vector ALU bound; the binary32 pair arithmetic runs without packed FP32, whose "
                               "peak then equals the FP64 one")
    return out


def qmdff_ethanol_box(cam, steps=200, nb=16):
    """The reference's examples/dynamic/ethanol_box on the GPU: 125 ethanol
    molecules (1125 atoms) in the periodic 27 angstrom box with the
    reference's own QMDFF (every molecule with the parameters of
    examples/poly_qmdff/ethanol/ethanol.qmdff), 16 beads, the Nose-Hoover
    chain at 200 K, dt 0.5 fs, the periodic wrap and transrot every step, as
    sampling.key runs it; FP64 and mixed precision."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import qmdff_fixtures as qf
    from caracal_amd import ff as cff
    p, x = qf.ethanol_box()
    amu = {1: 1.008, 6: 12.011, 8: 15.999}
    m = cam.mass_au([amu[int(z)] for z in p["at"]])
    beta = float(cam.beta_of(200.0))
    dt = 0.5 * cam.FS_TO_AU
    rng = np.random.default_rng(3)
    q = x[None] + rng.normal(0.0, 0.02, (nb,) + x.shape)
    pm = rng.normal(size=q.shape) * np.sqrt(m[None, :, None] * nb / beta)
    ff = cff.ForceField.qmdff(p)
    out = {"atoms": int(p["natoms"]), "nbeads": nb, "steps": steps}
    for name, prec in (("fp64", cam.CAM_PREC_FP64), ("mixed", cam.CAM_PREC_MIXED)):
        eng = cff.BoxEngine(ff, nb, dt, m, beta, None, precision=prec, nm_mode=cam.CAM_NM_REFERENCE)
        eng.set_nhc(20 * dt)
        eng.set_transrot(True)
        eng.set_state(q, pm)
        eng.prepare(steps)
        eng.run(10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms = eng.run(steps)
        wall = time.perf_counter() - t0
        eng.close()
        out[name] = {"bead_force_evals_per_s": nb * steps / wall, "ms_per_step": wall / steps * 1e3, "ms_gpu": ms}
    ff.close()
    return out


def force_matching(cam):
    """Force matching on the 3000-atom box with 64 beads: up to 50 L-BFGS
    iterations from the reference parameters perturbed by 10% (seed 3),
    gradients from the adjoint kernel."""
    import torch
    from caracal_amd import ff as cff, ffbox, fit
    s, xb = fit.default_problem(nbeads=64)
    ff = cff.ForceField(s, rc_angstrom=10.0)
    x = torch.tensor(xb, device="cuda")
    th_ref = ffbox.theta_ref()
    res = fit.force_match(ff, x, th_ref, fit.perturbed_theta(th_ref, 0.10, seed=3), steps=50)
    ff.close()
    return {"loss_first": res["loss"][0], "loss_last": res["loss"][-1], "iterations": res["iterations"],
            "evaluations": res["evaluations"], "ms_per_evaluation": res["s_per_eval"] * 1e3,
            "max_rel_err_force_determined_start": res["rel_err_forces0"],
            "max_rel_err_force_determined_end": res["rel_err_forces"],
            "max_rel_err_dispersion_start": res["rel_err_dispersion0"],
            "max_rel_err_dispersion_end": res["rel_err_dispersion"],
            "rel_param_err_end": [float(v) for v in np.round(res["theta"] / th_ref - 1.0, 5)],
            "nbeads": 64, "atoms": len(s["type"])}


def plain_fp64(cam, workloads, steps):
    """The plain form: one step per launch of a pair of trajectories, FP64,
    with the unfolded vector-ALU step (every bead transformed and evaluated)."""
    e, _ = prepare_children(cam, workloads, 2, 32, cam.CAM_PREC_FP64, steps_per_launch=1,
                            nm_kernel=cam.CAM_NMK_VALU)
    e.prepare(1)
    e.run(5)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.run(1, sync=False)
    e.sync()
    wall = time.perf_counter() - t0
    e.close()
    return 2 * 32 * steps / wall


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(steps=120000):
    """The CPU path timed on this host before the GPU is touched: the
    reference's own recrossing-child step (src/verlet.f90 with constrain = 2:
    FFTW normal modes, calc_xi of the centroid, egrad_ch4h for every bead),
    compiled from its sources at the reference's -O1 (and at -O2) by
    oracle/Makefile.ref, one 32-bead CH4H child per process on every core of
    this process's share (at most 16), as the reference's MPI task farm runs
    one trajectory per rank.  value = total bead-force-evals/s.  Falls back
    to the oracle's C port of the same step when the compiled reference is
    absent (kind "port")."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refstep
    from caracal_amd import workloads
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = min(16, ncpu)
    w = workloads.ch4h_children(ntraj=1, nbeads=32)
    rng = np.random.default_rng(7)
    p = workloads.mb_momenta(rng, w["masses"], 32, w["beta"][0], (32, 6, 3))
    res = {"unit": "bead-force-evals/s", "cores": cores, "cpu_model": cpu_model()}
    if not refstep.available("O1"):
        res.update(cpu_baseline_port(cores, 10.0))
        return res

    def farm(opt):
        with tempfile.TemporaryDirectory() as tmp:
            fin = os.path.join(tmp, "in.bin")
            refstep._write_input(fin, "ch4h", w["q"][0], p, w["masses"], w["dt"], w["beta"][0], steps, 2, 0,
                                 w["reaction"])
            t0 = time.perf_counter()
            procs = [subprocess.Popen([refstep.binary(opt), fin, os.path.join(tmp, f"o{i}.bin")],
                                      stdout=subprocess.PIPE, text=True, cwd=tmp) for i in range(cores)]
            secs = []
            for pr in procs:
                so, _ = pr.communicate()
                if pr.returncode != 0:
                    raise RuntimeError(f"reference step harness failed ({pr.returncode})")
                secs.append(float(so.split("seconds")[-1].split()[0]))
            wall = time.perf_counter() - t0
        return cores * steps * 32 / max(secs), max(secs), wall

    v1, s1, wall1 = farm("O1")
    res.update(value=v1, kind="reference",
               sample=f"{cores} processes x {steps} child steps of one 32-bead CH4H child trajectory "
                      f"(the reference's verlet + calc_xi + egrad_ch4h, amdflang -O1 as src/Makefile:44), "
                      f"slowest process {s1:.1f} s of step loop ({wall1:.1f} s wall with start-up)")
    if refstep.available("O2"):
        v2, s2, _ = farm("O2")
        res["O2"] = {"value": v2, "seconds": s2}
    return res


def cpu_baseline_port(cores, seconds):
    """The oracle's C port of the same child step, one trajectory per process."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(cores) as pool:
        out = pool.map(_cpu_worker, [(i, seconds) for i in range(cores)])
    steps = sum(r[0] for r in out)
    wall = max(r[1] for r in out)
    return dict(value=steps * 32 / wall, kind="port",
                sample=f"{cores} processes x {int(steps / cores)} child steps of one 32-bead CH4H trajectory "
                       f"(oracle C port of verlet + egrad_ch4h, gcc -O2), {wall:.1f} s")


def _cpu_worker(args):
    idx, seconds = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from caracal_amd import workloads
    w = workloads.ch4h_children(ntraj=1, nbeads=32, seed=idx)
    s = ol.make_sys(2, 32, w["dt"], w["masses"])
    rx = ol.make_rx(w["reaction"])
    rng = np.random.default_rng(idx)
    p = rng.normal(size=w["q"][0].shape) * 20
    t = ol.OracleTraj(s, w["q"][0], p, float(w["beta"][0]), rx=rx, ensemble=2)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            t.step()
        n += 10
    return n, time.perf_counter() - t0


def _surface_batch(cam, workloads, surface, w, ts, ntraj, nb, steps):
    """NVE trajectories of one analytic surface around its saddle (positions
    + N(0, 0.02^2) bohr, Maxwell-Boltzmann momenta), FP64, autotuned."""
    import torch
    rng = np.random.default_rng(2)
    na = ts.shape[0]
    q = ts[None, None] + rng.normal(0.0, 0.02, (ntraj, nb, na, 3))
    p = workloads.mb_momenta(rng, w["masses"], nb, w["beta"], (ntraj, nb, na, 3))
    e = cam.RPMDEngine(surface, nb, ntraj, w["dt"], w["masses"])
    e.set_state(q, p, np.full(ntraj, w["beta"]))
    _, best = e.autotune(50)
    e.prepare(steps)
    e.run(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.run(steps)
    wall = time.perf_counter() - t0
    e.close()
    return {"bead_force_evals_per_s": ntraj * nb * steps / wall, "us_per_step": wall / steps * 1e6,
            "ntraj": ntraj, "nbeads": nb, "steps": steps, "schedule": best}


def oh3_batch(cam, workloads, ntraj=4096, nb=32, steps=500):
    """OH + H2 (src/egrad_oh3.f): 4096 NVE trajectories of 32 beads at 300 K
    around the saddle, FP64, autotuned schedule; bead-force-evals/s."""
    return _surface_batch(cam, workloads, cam.CAM_SURFACE_OH3, workloads.oh3_nve(nbeads=nb), workloads.OH3_TS,
                          ntraj, nb, steps)


def brh2_batch(cam, workloads, ntraj=4096, nb=32, steps=500):
    """H + HBr on the BrH2 DIM-3C surface (src/egrad_brh2.f): 4096 NVE
    trajectories of 32 beads at 300 K around the saddle, FP64, autotuned
    schedule; bead-force-evals/s."""
    return _surface_batch(cam, workloads, cam.CAM_SURFACE_BRH2, workloads.brh2_nve(nbeads=nb), workloads.BRH2_TS,
                          ntraj, nb, steps)


def o3_batch(cam, workloads, ntraj=4096, nb=32, steps=500):
    """O + O2 on the Varga-Paukku-Truhlar O3 surface (src/egrad_o3.f): 4096
    NVE trajectories of 32 beads at 300 K around the exchange saddle, FP64,
    autotuned schedule; bead-force-evals/s."""
    return _surface_batch(cam, workloads, cam.CAM_SURFACE_O3, workloads.o3_nve(nbeads=nb), workloads.O3_TS,
                          ntraj, nb, steps)


def rpmd_rate_h_h2(scale=0.1):
    """The reference's H + H2 rate example (examples/calc_rate/h+h2/rate.key)
    end to end -- start structures, 110 umbrella windows, PMF, recrossing,
    k(T) -- with every step count times scale; wall time per phase."""
    from caracal_amd.rpmdrate import h_h2_example
    out = h_h2_example(scale=scale)
    return {"scale": scale, "wall_s": out["wall_s"], "total_wall_s": sum(out["wall_s"].values()),
            "xi_barrier": out["xi_barrier"], "dW_kcal_mol": (out["pmf_max"] - out["pmf_min"]) * 627.5094743,
            "kappa": out["kappa"], "k_cm3_molecule_s": out["k_cm3_molecule_s"],
            "note": "the key file's own settings (scale 1) took 53 s on one MI355X"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--ntraj", type=int, default=4096)
    ap.add_argument("--nbeads", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=120000,
                    help="child steps per host core of the CPU baseline (0: skip it)")
    ap.add_argument("--skip-extras", action="store_true", help="only the main FP64 batched measurement")
    ap.add_argument("--schedule", default=None,
                    help="fixed schedule 'valu|mfma|folded,STEPS_PER_LAUNCH[,WAVES_PER_EU]' (fused) or "
                         "'valu|mfma,split' instead of the autotuner (for profiling: every launch then has the same "
                         "shape)")
    ap.add_argument("--only", default=None, choices=["qmdff_box", "box", "ragged", "oh3", "brh2", "o3", "force_matching"],
                    help="run only this extra measurement and print its result")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB mebibytes five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); trajectories shard across GPUs "
              "without communication, see the sharding section of the design notes", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline runs first, before anything initialises the GPU, so its
    # worker processes start from a process without GPU state
    cpu = cpu_baseline(args.cpu_steps) if args.cpu_steps > 0 and not args.calibrate_copy else None

    import torch
    import caracal_amd as cam
    from caracal_amd import workloads
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 8 * n, "copies": 5}))
        return

    if args.only:
        fn = {"qmdff_box": lambda: qmdff_ethanol_box(cam), "box": lambda: box_nvt(cam),
              "ragged": lambda: ragged_batch(cam, workloads), "oh3": lambda: oh3_batch(cam, workloads),
              "brh2": lambda: brh2_batch(cam, workloads), "o3": lambda: o3_batch(cam, workloads),
              "force_matching": lambda: force_matching(cam)}[args.only]
        print(json.dumps({args.only: fn()}))
        return

    ntraj, nb = args.ntraj, args.nbeads
    if args.schedule:
        parts = args.schedule.split(",")
        unit, spl_fixed = parts[0], parts[1]
        wpe = int(parts[2]) if len(parts) > 2 else 1
        nmk = {"mfma": cam.CAM_NMK_MFMA, "folded": cam.CAM_NMK_FOLDED}.get(unit, cam.CAM_NMK_VALU)
        split = spl_fixed == "split"
        e, w = prepare_children(cam, workloads, ntraj, nb, cam.CAM_PREC_FP64, steps_per_launch=0 if split else int(spl_fixed),
                                nm_kernel=nmk, schedule=cam.CAM_SCHED_SPLIT if split else cam.CAM_SCHED_FUSED,
                                waves_per_eu=wpe)
        trials, best = [], {"nm_kernel": nmk, "split": int(split), "waves_per_eu": wpe,
                            "steps_per_launch": 1 if split else int(spl_fixed)}
    else:
        e, w = prepare_children(cam, workloads, ntraj, nb, cam.CAM_PREC_FP64)
        trials, best = e.autotune(100)
    wall, ms_launch, spl, nl = timed_run(e, args.steps, args.warmup)
    num, den = e.kappa(args.steps)
    e.close()
    bead_evals = ntraj * nb * args.steps
    value = bead_evals / wall
    steps_per_s = ntraj * args.steps / wall
    folded = best["nm_kernel"] == cam.CAM_NMK_FOLDED
    # algorithmic work of the timed run over its GPU time (events around the
    # graph replays): nl launches of ms_launch each cover exactly args.steps
    flops_run = step_flops(6, nb, folded) * ntraj * args.steps
    achieved = flops_run / (ms_launch * nl * 1e-3) / 1e12
    if achieved > FP64_PEAK_TFLOPS:
        raise RuntimeError(f"roofline above peak ({achieved:.1f} TFLOP/s): the work count does not match the run")
    nmk_name = NMK_NAMES[best["nm_kernel"]]
    if best.get("split"):
        kernel = (f"split step: step_kernel<SurfCH4H, {nb}, double, CHILD, {nmk_name.upper()} normal modes, "
                  f"no forces> + force_kernel<SurfCH4H, {nb}, double>")
    elif folded:
        kernel = f"fold_step_kernel<SurfCH4H, {nb}, double, CHILD, {best.get('waves_per_eu', 1)}>"
    else:
        kernel = f"step_kernel<SurfCH4H, {nb}, double, CHILD, {nmk_name.upper()} normal modes>"
    distinct = nb // 2 + 1 if folded else nb

    out = {
        "metric": "bead-force-evals/s (batched RPMD children, CH4H, 32 beads x 4096 trajectories)",
        "value": value,
        "unit": "bead-force-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded saddle perturbations, Maxwell-Boltzmann momenta)",
        "config": {"workload": "ch4h_children", "surface": "CH5 (egrad_ch4h)", "ntraj": ntraj, "nbeads": nb,
                   "natoms": 6, "dt_fs": 0.1, "kelvin": 500, "precision": "fp64",
                   "nm_mode": "reference cosine transform", "steps_per_launch": spl,
                   "nm_kernel": nmk_name, "schedule": "split" if best.get("split") else "fused"},
        "rpmd_steps_per_s": steps_per_s,
        "distinct_bead_force_evals_per_s": ntraj * distinct * args.steps / wall,
        "value_note": ("value counts the forces of all nbeads beads of every trajectory step, as the reference "
                       "evaluates them; the reference transform keeps beads j and nbeads - j equal, so the folded "
                       f"kernel evaluates {distinct} distinct beads per step and mirrors the rest"),
        "kappa_final": float(num[-1] / den) if den > 0 else None,
        "autotune": trials,
        "roofline": {"bound": "valu-fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": launch_traffic(nmk_name, ntraj, nb, spl),
                     "kernel": kernel, "ms_per_launch": ms_launch, "launches": nl, "steps_per_launch": spl,
                     "flops_per_launch": flops_run / nl,
                     "flops_split_per_trajectory_step": step_flops_parts(6, nb, folded),
                     "note": ("FP64 peak is the same for the vector ALUs and the matrix cores on MI355X; "
                              f"{100 * step_flops_parts(6, nb, folded)['surface'] / step_flops(6, nb, folded):.0f}% "
                              "of the counted flops are the surface on the vector ALUs; elementary functions not "
                              "counted; traffic is the counter-measured HBM bytes of one launch of this shape")},
    }
    if not args.skip_extras:
        e, _ = prepare_children(cam, workloads, ntraj, nb, cam.CAM_PREC_MIXED)
        mtrials, mbest = e.autotune(100)
        wall_m, ms_m, spl_m, _ = timed_run(e, args.steps, args.warmup)
        e.close()
        mixed_value = bead_evals / wall_m
        plain = plain_fp64(cam, workloads, 200)
        ferr, ferr_bead, ferr_floor = force_error_mixed(cam, workloads, 256, nb)
        out["mixed"] = {"value": mixed_value, "rpmd_steps_per_s": ntraj * args.steps / wall_m,
                        "dtype": "fp32 forces / fp64 state", "rel_force_err": ferr,
                        "rel_force_err_worst_bead": ferr_bead, "rel_force_err_worst_bead_floor_1e-3": ferr_floor,
                        "schedule": mbest}
        out["energy_drift"] = energy_drift(cam, workloads)
        out["plain_fp64_one_traj_per_launch"] = {"value": plain}
        out["speedup_mixed_vs_plain"] = mixed_value / plain
        out["speedup_fp64_batched_vs_plain"] = value / plain
        out["ragged"] = ragged_batch(cam, workloads)
        out["oh3"] = oh3_batch(cam, workloads)
        out["brh2"] = brh2_batch(cam, workloads)
        out["o3"] = o3_batch(cam, workloads)
        out["box"] = box_nvt(cam)
        out["qmdff_ethanol_box"] = qmdff_ethanol_box(cam)
        out["force_matching"] = force_matching(cam)
        out["rpmd_rate_h_h2"] = rpmd_rate_h_h2()
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
