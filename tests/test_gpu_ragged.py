"""Ragged batches through the native batch layer: buckets of different
surfaces and bead counts run concurrently and give exactly what each bucket
gives on its own; mixed precision stays within 1e-5 of FP64 per bucket."""
import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd import workloads
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _rel_err(a, b):
    num = np.abs(a - b).reshape(a.shape[0] * a.shape[1], -1).max(axis=1)
    den = np.abs(b).reshape(b.shape[0] * b.shape[1], -1).max(axis=1)
    return num.max() / den.max()


def test_ragged_batch_equals_buckets_run_alone():
    buckets = workloads.ragged(ntraj=600, seed=1)
    assert len(buckets) == 10
    rb = cam.RaggedBatch(buckets, precision=cam.CAM_PREC_MIXED, steps_per_launch=10)
    rb.set_states()
    ms = rb.run(30)
    print(f"ragged batch of {sum(len(d['beta']) for _, _, d in buckets)} trajectories: 30 steps in {ms:.3f} ms")
    for (surf, nb, d), eb in zip(buckets, rb.engines):
        alone = cam.RPMDEngine(surf, nb, len(d["beta"]), d["dt"], d["masses"], precision=cam.CAM_PREC_MIXED,
                               steps_per_launch=10)
        alone.set_state(d["q"], d["p"], d["beta"])
        alone.run(30)
        qa, pa = alone.get_state()
        qb, pb = eb.get_state()
        alone.close()
        assert np.array_equal(qa, qb) and np.array_equal(pa, pb), (surf, nb)
    rb.close()


def test_ragged_mixed_forces_within_1e5_of_fp64_per_bucket():
    buckets = workloads.ragged(ntraj=2000, seed=1)
    errs = {}
    g = {}
    for prec in (cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED):
        rb = cam.RaggedBatch(buckets, precision=prec)
        rb.set_states()
        g[prec] = [e.get_forces()[1] for e in rb.engines]
        rb.close()
    for i, (surf, nb, _) in enumerate(buckets):
        errs[(surf, nb)] = _rel_err(g[cam.CAM_PREC_MIXED][i], g[cam.CAM_PREC_FP64][i])
    print("mixed vs fp64 relative force error per (surface, nbeads):",
          {f"{'H3' if s == cam.CAM_SURFACE_H3 else 'CH4H'}/{n}": f"{v:.2e}" for (s, n), v in errs.items()})
    assert max(errs.values()) <= 1e-5


def test_ragged_batch_with_five_surfaces_in_fp64():
    """H3, CH5, OH + H2, BrH2 and O3 buckets (4-32 beads) in one FP64 batch,
    each equal bit for bit to its bucket run alone (the batch's folded
    kernels take two LDS column buffers, an engine alone one)."""
    rng = np.random.default_rng(4)
    systems = {cam.CAM_SURFACE_OH3: (workloads.oh3_nve, workloads.OH3_TS),
               cam.CAM_SURFACE_H3: (workloads.h3_nve, workloads.H3_TS),
               cam.CAM_SURFACE_BRH2: (workloads.brh2_nve, workloads.BRH2_TS),
               cam.CAM_SURFACE_O3: (workloads.o3_nve, workloads.O3_TS)}
    buckets = []
    for surf, nb, n in ((cam.CAM_SURFACE_OH3, 16, 40), (cam.CAM_SURFACE_H3, 8, 30), (cam.CAM_SURFACE_CH4H, 4, 20),
                        (cam.CAM_SURFACE_OH3, 32, 10), (cam.CAM_SURFACE_BRH2, 32, 20), (cam.CAM_SURFACE_O3, 16, 30),
                        (cam.CAM_SURFACE_BRH2, 6, 25)):
        if surf in systems:
            w = systems[surf][0](nbeads=nb)
            ts = systems[surf][1]
        else:
            w = workloads.ch4h_children(ntraj=1, nbeads=nb)
            ts = workloads.CH4H_TS
        beta = np.full(n, float(np.atleast_1d(w["beta"])[0]))
        q = ts[None, None] + rng.normal(0.0, 0.02, (n, nb) + ts.shape)
        p = workloads.mb_momenta(rng, w["masses"], nb, beta[0], (n, nb) + ts.shape)
        buckets.append((surf, nb, dict(q=q, p=p, beta=beta, masses=w["masses"], dt=w["dt"])))
    rb = cam.RaggedBatch(buckets, steps_per_launch=10)
    rb.set_states()
    rb.run(30)
    for (surf, nb, d), eb in zip(buckets, rb.engines):
        alone = cam.RPMDEngine(surf, nb, len(d["beta"]), d["dt"], d["masses"], steps_per_launch=10)
        alone.set_state(d["q"], d["p"], d["beta"])
        alone.run(30)
        qa, pa = alone.get_state()
        qb, pb = eb.get_state()
        alone.close()
        assert np.array_equal(qa, qb) and np.array_equal(pa, pb), (surf, nb)
    rb.close()


def test_ragged_8192_mixed_forces_and_fp64_steps_match_oracle():
    """The ragged batch at its size (8192 trajectories, H3 and CH5 at 4-64
    beads, 200-1000 K).  Mixed precision: every bucket's start forces on 16
    sampled trajectories (all their beads) within 1e-5 of the oracle's
    surfaces relative to the largest |F| of the bucket; per bead, relative to
    max(|F| of the bead, 1e-3 hartree/bohr), within 1e-4 (binary32 surface
    arithmetic on beads with small forces reaches ~2e-5).  FP64: 50 steps of
    the whole batch, two trajectories per bucket against the oracle's verlet
    to 1e-10."""
    buckets = workloads.ragged(ntraj=8192, seed=1)
    assert sum(len(d["beta"]) for _, _, d in buckets) == 8192
    rng = np.random.default_rng(5)
    rb = cam.RaggedBatch(buckets, precision=cam.CAM_PREC_MIXED)
    rb.set_states()
    worst = {}
    for (surf, nb, d), eng in zip(buckets, rb.engines):
        _, g = eng.get_forces()
        egrad = ol.egrad_h3 if surf == cam.CAM_SURFACE_H3 else ol.egrad_ch4h
        pick = rng.choice(len(d["beta"]), min(16, len(d["beta"])), replace=False)
        ref = np.stack([egrad(d["q"][t])[1] for t in pick])
        dev = np.abs(g[pick] - ref)
        whole = dev.max() / np.abs(ref).max()
        per_bead = (dev.reshape(len(pick), nb, -1).max(axis=2)
                    / np.maximum(np.abs(ref).reshape(len(pick), nb, -1).max(axis=2), 1e-3)).max()
        worst[(surf, nb)] = (whole, per_bead)
        assert whole <= 1e-5 and per_bead <= 1e-4, (surf, nb, whole, per_bead)
    rb.close()
    print("mixed vs oracle force error (bucket, per bead):",
          {f"{'H3' if s == cam.CAM_SURFACE_H3 else 'CH4H'}/{n}": f"{a:.1e}/{b:.1e}" for (s, n), (a, b) in worst.items()})
    rb = cam.RaggedBatch(buckets, precision=cam.CAM_PREC_FP64, steps_per_launch=25)
    rb.set_states()
    rb.run(50)
    for (surf, nb, d), eng in zip(buckets, rb.engines):
        q, p = eng.get_state()
        s = ol.make_sys(surf, nb, d["dt"], d["masses"])
        for t in (0, len(d["beta"]) - 1):
            tr = ol.OracleTraj(s, d["q"][t], d["p"][t], float(d["beta"][t]), ensemble=0)
            for _ in range(50):
                tr.step()
            assert np.abs(q[t] - tr.q).max() <= 1e-10 * np.abs(tr.q).max(), (surf, nb, t)
            assert np.abs(p[t] - tr.p).max() <= 1e-8 * np.abs(tr.p).max(), (surf, nb, t)
    rb.close()


def test_ragged_batch_graph_replays_and_recaptures():
    """The batch runs as one captured graph per replay length.  Runs of 10,
    25 and again 10 steps (two graphs, the first replayed) and a new state in
    between give bit for bit what each bucket gives run alone.  After the
    autotuner changed the buckets' schedules (each engine drops its graphs,
    so the batch captures again), a batch run equals each bucket's engine run
    alone from the same state."""
    buckets = workloads.ragged(ntraj=300, seed=2)
    rb = cam.RaggedBatch(buckets, precision=cam.CAM_PREC_FP64, steps_per_launch=5)
    rb.set_states()
    alone = []
    for surf, nb, d in buckets:
        e = cam.RPMDEngine(surf, nb, len(d["beta"]), d["dt"], d["masses"], precision=cam.CAM_PREC_FP64,
                           steps_per_launch=5)
        e.set_state(d["q"], d["p"], d["beta"])
        alone.append(e)

    def run(n):
        rb.run(n)
        for e in alone:
            e.run(n)

    def same():
        for (surf, nb, _), eb, ea in zip(buckets, rb.engines, alone):
            qa, pa = ea.get_state()
            qb, pb = eb.get_state()
            assert np.array_equal(qa, qb) and np.array_equal(pa, pb), (surf, nb)

    run(10)
    run(25)
    run(10)
    same()
    for engs in (rb.engines, alone):                 # momenta reversed
        for e, (_, _, d) in zip(engs, buckets):
            q, p = e.get_state()
            e.set_state(q, -p, d["beta"])
    run(10)
    same()
    for e in alone:
        e.close()
    rb.autotune(10)
    start = [e.get_state() for e in rb.engines]
    # the state includes the forces: both runs start from set_state's
    for (_, _, d), e, (q0, p0) in zip(buckets, rb.engines, start):
        e.set_state(q0, p0, d["beta"])
    rb.run(20)
    after = [e.get_state() for e in rb.engines]
    for (surf, nb, d), e, (q0, p0), (q1, p1) in zip(buckets, rb.engines, start, after):
        e.set_state(q0, p0, d["beta"])
        e.run(20)
        q, p = e.get_state()
        assert np.array_equal(q, q1) and np.array_equal(p, p1), (surf, nb)
    rb.close()
