"""The oracle's Nose-Hoover chain (src/nhc.f90:35-169), pinned by the
conserved energy of the extended system: the ring polymer's Hamiltonian over
nbeads^2 (the chain acts on the centroid-momenta kinetic energy, which is the
ring polymer's over nbeads^2), plus the thermostats' kinetic energies
qnh v^2/2 and their potential nfree kT eta_1 + kT (eta_2 + eta_3 + eta_4).
Second order in dt, and the thermostat does act (the ring polymer's own
energy does not stay constant)."""
import numpy as np

import oracle_lib as ol
from caracal_amd import workloads
from test_oracle_rpmd import _hamiltonian


def _run(dt_scale, nsteps, nose_steps=20.0):
    w = workloads.h3_nve(nbeads=8)
    rev = (-np.arange(8)) % 8
    q = 0.5 * (w["q"] + w["q"][rev])
    p = 0.5 * (w["p"] + w["p"][rev]) * 1.7          # hot start: the chain must pull energy out
    s = ol.make_sys(1, 8, w["dt"] * dt_scale, w["masses"])
    t = ol.OracleTraj(s, q, p, w["beta"], ensemble=5)
    t.nhc_begin(nose_steps * w["dt"])
    kelvin = 1.0 / (w["beta"] * 0.316679e-5)
    ekt = float(np.float32(np.float32(1.380649e-23) / np.float32(4.3597447e-18))) * kelvin
    nfree = 9

    def energies():
        h = _hamiltonian(s, t.q, t.p, t.v, w["beta"]) / 64.0
        ext = np.sum(0.5 * t.qnh * t.vnh ** 2) + nfree * ekt * t.eta[0] + ekt * np.sum(t.eta[1:])
        return h, h + ext

    h0, e0 = energies()
    hs, es = [], []
    for _ in range(nsteps):
        t.step()
        h, e = energies()
        hs.append(h - h0)
        es.append(e - e0)
    return np.array(hs), np.array(es)


def test_nhc_conserves_the_extended_energy_to_second_order():
    h1, e1 = _run(1.0, 400)
    h2, e2 = _run(0.5, 800)
    assert np.max(np.abs(h1)) > 20 * np.max(np.abs(e1))          # the thermostat moves energy
    ratio = np.max(np.abs(e1)) / np.max(np.abs(e2))
    assert 3.0 < ratio < 5.0
