"""Writes tests/golden/wham_reference.npz: inputs of calc_rate's WHAM call
(src/calc_rate.f90:2012) and the PMF the reference's own src/wham.f90
returns for them, through the compiled reference (oracle/_ref/ref_wham,
built by oracle/Makefile.ref from the reference's sources).

Two cases: 16 windows of multinomial samples (20,000 each) of a double-well
PMF under harmonic window biases on 401 bins; and 9 windows over a range
twice as wide as they cover, so that bins without any sample take the
routine's 1e-15 floor.  Every window holds samples.

    python tests/golden/make_wham_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_wham")
KB = 0.316679e-5


def case(seed, nwin, nbins, xi_min, xi_max, w_lo, w_hi, k_force, kelvin, nsamp):
    rng = np.random.default_rng(seed)
    bs = (xi_max - xi_min) / (nbins - 1)
    x = xi_min + bs * np.arange(1, nbins + 1)
    pmf = 0.006 * (x * x - 0.25) ** 2 / 0.0625 + 0.001 * x      # hartree: wells at -0.5 and 0.5, tilted
    beta = 1.0 / (KB * kelvin)
    xi_wins = np.linspace(w_lo, w_hi, nwin)
    hist = np.empty((nwin, nbins))
    for i, xw in enumerate(xi_wins):
        p = np.exp(-beta * (pmf + k_force * (x - xw) ** 2))
        hist[i] = rng.multinomial(nsamp, p / p.sum())
    return dict(nbins=nbins, xi_min=xi_min, xi_max=xi_max, kelvin=kelvin, xi_wins=xi_wins,
                k_force=np.full(nwin, k_force), hist=hist)


def run_reference(c):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([c["nbins"], len(c["xi_wins"])], dtype=np.int32).tofile(f)
            np.array([c["xi_min"], c["xi_max"], c["kelvin"]], dtype=np.float64).tofile(f)
            c["xi_wins"].astype(np.float64).tofile(f)
            c["k_force"].astype(np.float64).tofile(f)
            np.asfortranarray(c["hist"]).T.astype(np.float64).tofile(f)     # column-major hist(i, j)
        subprocess.run([REF, fin, fout], check=True, cwd=d, capture_output=True)
        return np.fromfile(fout, dtype=np.float64)


def main():
    cases = [case(3, 16, 401, -0.7, 0.7, -0.6, 0.6, 0.08, 300.0, 20000),
             case(4, 9, 301, -1.2, 1.2, -0.5, 0.5, 0.12, 500.0, 50000)]
    out = {}
    for n, c in enumerate(cases):
        for k, v in c.items():
            out[f"c{n}_{k}"] = np.asarray(v)
        out[f"c{n}_pmf_ref"] = run_reference(c)
    np.savez_compressed(os.path.join(HERE, "wham_reference.npz"), ncases=len(cases), **out)


if __name__ == "__main__":
    sys.exit(main())
