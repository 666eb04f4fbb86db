"""The oracle's water surface (oracle/water.c) against the compiled
reference (src/egrad_water.f90 through oracle/ref_water_harness.f90), central
differences of its own energy where the reference's gradient is the
derivative of its energy (plain Coulomb), and closed forms."""
import numpy as np
import pytest

import geometries as geo
import oracle_lib as ol

BOHR = 0.52917721092


def _zahn(rc_angstrom):
    from math import erfc, exp, pi, sqrt
    za = 0.2 * BOHR
    rc = rc_angstrom / BOHR
    zp = erfc(za * rc) / rc ** 2 + 2 * za / sqrt(pi) * exp(-(za * rc) ** 2) / rc
    return za, zp, rc


CASES = {
    "cluster": dict(periodic=False),
    "box_plain": dict(periodic=True, box=18.0 / BOHR, cut=8.0 / BOHR),
    "box_zahn": dict(periodic=True, box=18.0 / BOHR, zahn=True),
}


def _case(name):
    kw = dict(CASES[name])
    if kw.get("zahn"):
        za, zp, rc = _zahn(8.0)
        kw.update(zahn_a=za, zahn_par=zp, cut=rc)
    return kw


@pytest.mark.parametrize("name", list(CASES))
def test_water_oracle_matches_compiled_reference(name):
    if ol.ref() is None:
        pytest.skip("compiled reference not built here")
    kw = _case(name)
    box = kw.get("box", 12.0 / BOHR)
    x = geo.water_box(64 if kw.get("periodic") else 8, box, 11)
    e, g = ol.water_eval(x, **kw)
    er, gr = ol.ref_water_eval(x, **kw)
    assert abs(e - er) <= 1e-12 * abs(er)
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-12 * np.abs(gr).max())


def test_water_gradient_is_the_derivative_with_plain_coulomb():
    x = geo.water_box(6, 10.0 / BOHR, 12)
    _, g = ol.water_eval(x)
    h = 1e-5
    rng = np.random.default_rng(0)
    for idx in rng.choice(x.size, 25, replace=False):
        xp, xm = x.copy().ravel(), x.copy().ravel()
        xp[idx] += h
        xm[idx] -= h
        fd = (ol.water_eval(xp.reshape(x.shape))[0] - ol.water_eval(xm.reshape(x.shape))[0]) / (2 * h)
        assert abs(fd - g.ravel()[idx]) < 1e-7 + 1e-6 * abs(fd)


def test_single_molecule_closed_forms():
    p = ol.water_params()
    r0, r0hh, de, a, khh, krt, krr = p[:7]
    # equilibrium geometry: every term vanishes
    half = np.arcsin(0.5 * r0hh / r0)
    x = np.array([[0, 0, 0], [r0 * np.sin(half), r0 * np.cos(half), 0], [-r0 * np.sin(half), r0 * np.cos(half), 0]])
    e, g = ol.water_eval(x)
    assert abs(e) < 1e-15 and np.abs(g).max() < 1e-13
    # one O-H stretched by d along its bond: Morse with exp(+a d), coupling terms
    d = 0.05
    x2 = x.copy()
    x2[1] *= (r0 + d) / r0
    e2, _ = ol.water_eval(x2)
    rhh = np.linalg.norm(x2[1] - x2[2])
    dh = rhh - r0hh
    want = de * (1 - np.exp(a * d)) ** 2 + 0.5 * khh * dh ** 2 + krt * dh * d
    assert abs(e2 - want) < 1e-14 * max(1.0, abs(want))


def test_lj_applies_when_the_lower_index_is_oxygen():
    """Two molecules far apart on the x axis, one atom pair at a time: the
    LJ term enters for (O of molecule 1, any atom of molecule 2) and for no
    pair whose lower-index atom is a hydrogen."""
    p = ol.water_params()
    qh, qo, sig, eps = p[7], p[8], p[9], p[10]
    half = np.arcsin(0.5 * p[1] / p[0])
    mol = np.array([[0, 0, 0], [p[0] * np.sin(half), p[0] * np.cos(half), 0],
                    [-p[0] * np.sin(half), p[0] * np.cos(half), 0]])
    x = np.vstack([mol, mol + np.array([7.0, 0.3, -0.2])])
    e, _ = ol.water_eval(x)
    want = 0.0
    q = [qo, qh, qh, qo, qh, qh]
    for i in range(3):
        for j in range(3, 6):
            r = np.linalg.norm(x[i] - x[j])
            want += q[i] * q[j] / r
            if i == 0:
                s6 = (sig / r) ** 6
                want += 4 * eps * (s6 * s6 - s6)
    assert abs(e - want) < 1e-13 * abs(want)
