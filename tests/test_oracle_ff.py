"""Pins the oracle's QMDFF-form box: central differences of the energy in the
positions (gradnum, src/gradnum.f90:36-241) and in the parameters, of the
force-matching contraction v . dE/dx in the parameters, and closed forms of
isolated terms."""
import math

import numpy as np

import oracle_lib as ol
from caracal_amd import ffbox


def small_box(nmol=27, seed=4):
    s = ffbox.box_system(nmol=nmol, box_angstrom=9.0, seed=seed)
    s["x"] = s["x"] + np.random.default_rng(seed).normal(0, 0.05, s["x"].shape)
    return s


def test_forces_are_central_differences_of_energy():
    s = small_box()
    ff = ol.FFOracle(s, rc_angstrom=4.4)
    th = ffbox.theta_ref()
    e, g, _, _ = ff.eval(th, s["x"])
    h = 1e-5
    rng = np.random.default_rng(0)
    for idx in rng.choice(s["x"].size, 24, replace=False):
        xp, xm = s["x"].copy().ravel(), s["x"].copy().ravel()
        xp[idx] += h
        xm[idx] -= h
        fd = (ff.eval(th, xp.reshape(-1, 3), want_grad=False)[0] - ff.eval(th, xm.reshape(-1, 3), want_grad=False)[0]) / (2 * h)
        assert abs(fd - g.ravel()[idx]) < 1e-7 + 1e-6 * abs(fd)


def test_parameter_gradient_is_central_difference():
    s = small_box()
    ff = ol.FFOracle(s, rc_angstrom=4.4)
    th = ffbox.theta_ref()
    v = np.random.default_rng(1).normal(size=s["x"].shape)
    _, g, dth, dvth = ff.eval(th, s["x"], v=v)
    for p in range(10):
        h = 1e-6 * max(1.0, abs(th[p]))
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        ep, gp, _, _ = ff.eval(tp, s["x"])
        em, gm, _, _ = ff.eval(tm, s["x"])
        assert abs((ep - em) / (2 * h) - dth[p]) < 1e-6 * max(1.0, abs(dth[p])), p
        fdv = (np.sum(v * gp) - np.sum(v * gm)) / (2 * h)
        assert abs(fdv - dvth[p]) < 1e-6 * max(1.0, abs(dvth[p])), p


def test_isolated_terms_closed_forms():
    th = ffbox.theta_ref()
    # one molecule at equilibrium: bonded energy zero, no nonbonded partner
    s = ffbox.box_system(nmol=1, box_angstrom=30.0, seed=0)
    ff = ol.FFOracle(s)
    e, g, _, _ = ff.eval(th, s["x"])
    assert abs(e) < 1e-24 and np.abs(g).max() < 1e-12
    # two molecules far apart except one O-O pair at distance d: energy is
    # q_O^2 phi(d) + C6 disp + rep for that pair, plus O-H and H-H pairs...
    # check the Coulomb-only closed form with a single pair of charges by
    # switching off dispersion (C6 = 0) and repulsion is fixed -> subtract it
    two = dict(x=np.array([[1.0, 1.0, 1.0], [5.0, 1.0, 1.0]]), type=np.array([0, 0], np.int32),
               mol=np.array([0, 1], np.int32), bonds=np.zeros((0, 2), np.int32), angles=np.zeros((0, 3), np.int32),
               box=40.0)
    ff2 = ol.FFOracle(two)
    t2 = th.copy()
    t2[7] = 0.0
    e2, _, _, _ = ff2.eval(t2, two["x"])
    a, rc, zp = ffbox.zahn_params()
    d = 4.0
    coul = th[5] ** 2 * (math.erfc(a * d) / d - zp * (d - rc))
    rep = 36.0 * math.exp(-1.5 * d) / d
    assert abs(e2 - (coul + rep)) < 1e-14
    # dispersion closed form
    t3 = th.copy()
    t3[5] = 0.0
    e3, _, _, _ = ff2.eval(t3, two["x"])
    sr42, r0, _, _ = ffbox.pair_tables()
    disp = -th[7] * (1 / (d ** 6 + r0[0, 0] ** 6) + sr42[0, 0] / (d ** 8 + r0[0, 0] ** 8))
    assert abs(e3 - (disp + rep)) < 1e-15
    # minimum image: shifting one atom by a box vector changes nothing
    x = two["x"].copy()
    x[1, 0] += 40.0
    assert ff2.eval(th, x)[0] == ff2.eval(th, two["x"])[0]
