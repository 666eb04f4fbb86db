"""Runs the reference's own RPMD step (oracle/_ref/ref_step_O1, the compiled
src/verlet.f90 with its FFTW normal-mode transforms, transrot, calc_xi and
analytic surfaces, driven by oracle/ref_step_harness.f90).  Test
infrastructure only: used to pin the oracle's step and to write golden
trajectories, never on a product path."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
SURFACE_CODE = {"h3": 1, "ch4h": 2, "oh3": 3, "qmdff": 4, "brh2": 5, "o3": 6}


def binary(opt="O1"):
    return os.path.join(REF_DIR, f"ref_step_{opt}")


def available(opt="O1"):
    return os.path.exists(binary(opt))


def _write_input(path, surface, q, p, masses, dt, beta, nsteps, constrain, every, reaction, thermostat=0,
                 andersen_step=0, kelvin=None, nose_q=0.0, k_force=0.0, qmdff=None):
    nbeads, natoms, _ = q.shape
    with open(path, "wb") as f:
        np.array([natoms, nbeads, nsteps, constrain, SURFACE_CODE[surface], every, thermostat, andersen_step],
                 dtype=np.int32).tofile(f)
        np.array([dt, beta], dtype=np.float64).tofile(f)
        np.asarray(masses, dtype=np.float64).tofile(f)
        np.ascontiguousarray(q, dtype=np.float64).tofile(f)
        np.ascontiguousarray(p, dtype=np.float64).tofile(f)
        if thermostat == 2:
            np.array([kelvin, nose_q], dtype=np.float64).tofile(f)
        if surface == "qmdff":
            name = os.path.abspath(qmdff["file"]).encode()
            np.array([len(name)], dtype=np.int32).tofile(f)
            f.write(name)
            np.array([int(qmdff.get("periodic", False)), int(qmdff.get("zahn", False))], dtype=np.int32).tofile(f)
            np.array([qmdff.get("box", 0.0), qmdff.get("coul_cut", 50.0), qmdff.get("vdw_cut", 10.0)],
                     dtype=np.float64).tofile(f)
        if constrain >= 0:
            r = reaction
            np.array([len(r.bond_form), len(r.bond_break)], dtype=np.int32).tofile(f)
            np.asarray(r.bond_form, dtype=np.int32).reshape(-1).tofile(f)
            np.asarray(r.bond_break, dtype=np.int32).reshape(-1).tofile(f)
            np.asarray(r.form_ref, dtype=np.float64).tofile(f)
            np.asarray(r.break_ref, dtype=np.float64).tofile(f)
            counts = [len(a) for a in r.reactants]
            np.array([len(counts)], dtype=np.int32).tofile(f)
            np.asarray(counts, dtype=np.int32).tofile(f)
            maxr = max(counts)
            ra = np.zeros((len(counts), maxr), dtype=np.int32)
            for k, atoms in enumerate(r.reactants):
                ra[k, :len(atoms)] = atoms
            ra.tofile(f)
            np.array([r.r_inf, r.xi_ideal], dtype=np.float64).tofile(f)
        if constrain == 0:
            np.array([k_force], dtype=np.float64).tofile(f)


def run(surface, q, p, masses, dt, beta, nsteps, constrain=-1, every=1, reaction=None, opt="O1", **kw):
    """nsteps reference verlet steps from (q, p), each (nbeads, natoms, 3),
    after the reference's mdinit.  constrain -1: dynamic.x (transrot after
    every step); 0: umbrella window (k_force=..., the reaction's xi_ideal;
    bias in the start forces and every step, transrot); 2: recrossing child
    (xi of the centroid every step).  thermostat=1 (Andersen every
    andersen_step steps, the compiler's random stream) or 2 (Nose-Hoover
    chain at kelvin=..., nose_q=...).  Returns (frames, seconds): frames is a
    dict of q, p (nframes, nbeads, natoms, 3), epot and xi (nframes) after
    every 'every' steps.  surface "qmdff" takes qmdff=dict(file=..., periodic=...,
    box=..., coul_cut=..., vdw_cut=..., zahn=...) (angstrom), as the key file."""
    q = np.asarray(q, dtype=np.float64)
    nbeads, natoms, _ = q.shape
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        _write_input(fin, surface, q, p, masses, dt, beta, nsteps, constrain, every, reaction, **kw)
        out = subprocess.run([binary(opt), fin, fout], check=True, capture_output=True, text=True, cwd=tmp)
        seconds = float(out.stdout.split("seconds")[-1].split()[0])
        raw = np.fromfile(fout, dtype=np.float64) if every > 0 else np.zeros(0)
    n = 2 * nbeads * natoms * 3 + 2
    raw = raw.reshape(-1, n)
    m = nbeads * natoms * 3
    frames = dict(q=raw[:, :m].reshape(-1, nbeads, natoms, 3), p=raw[:, m:2 * m].reshape(-1, nbeads, natoms, 3),
                  epot=raw[:, 2 * m], xi=raw[:, 2 * m + 1])
    return frames, seconds
