"""Pins the oracle's OH + H2 surface (src/egrad_oh3.f): against the compiled
reference (oracle/_ref), against central differences of its energy, against
the H2 <-> H3 exchange symmetry, and against the closed form of the separated
fragments.  Far apart, the LEPS term over (R2, R3, R6) reduces to a Morse
curve of H2 with (DE3, BETA3, RE3) less DE3, the H2O and four-body terms
vanish, and with the offset -2 DE2 + DE3 the surface is
Morse(DE1, BETA1, RE1; r_OH) + Morse(DE3, BETA3, RE3; r_HH)."""
import numpy as np
import pytest

import oracle_lib as ol
from caracal_amd import workloads

needs_ref = pytest.mark.skipif(ol.ref() is None, reason="compiled reference not built")

DE = (0.148201, 0.0275690, 0.151548)
BETA = (1.260580, 0.924180, 1.068620)
RE = (1.863300, 2.907700, 1.428600)


def _cloud(n, sigma, seed):
    rng = np.random.default_rng(seed)
    return workloads.OH3_TS[None] + rng.normal(0.0, sigma, (n, 4, 3))


def _separated(r_oh, r_hh, dist):
    """OH along x at the origin, H2 along y, centres dist bohr apart on z."""
    return np.array([[0, 0, 0], [r_oh, 0, 0], [0, -r_hh / 2, dist], [0, r_hh / 2, dist]], dtype=np.float64)


def _special():
    out = [_separated(1.8633, 1.4286, 60.0), _separated(2.4, 1.2, 45.0)]
    t = workloads.OH3_TS
    # compact, water-like product side, and one O-H distance in each VH2O
    # switching branch (x < 15, 15 <= x < 43, x >= 43 -> about 14.3 and 37.6 bohr)
    out.append(np.array([[0, 0, 0], [1.8, 0, 0], [-0.5, 1.75, 0], [-0.9, 4.5, 0.3]]))
    out.append(np.array([[0, 0, 0], [1.1, 0.2, 0], [-0.3, 1.0, 0.2], [0.6, 1.2, -0.4]]))
    for d in (10.0, 20.0, 30.0, 40.0, 90.0):
        x = t.copy()
        x[2:] += np.array([d, 0.0, 0.0])
        out.append(x)
    # the VH2O quirk: O-H2 beyond 37.6 bohr with a stretched H2, so the stale
    # three-body derivative is not small
    out.append(np.array([[0, 0, 0], [1.9, 0, 0], [40.0, 0, 0], [42.2, 0.3, 0]]))
    return np.array(out)


@needs_ref
@pytest.mark.parametrize("q", [_cloud(300, 0.15, 0), _cloud(200, 0.5, 1), _special()],
                         ids=["near-saddle", "wide", "special"])
def test_oh3_matches_compiled_reference(q):
    v, g = ol.egrad_oh3(q)
    vr, gr = ol.ref_egrad_oh3(q)
    np.testing.assert_allclose(v, vr, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-13 * np.abs(gr).max())


def _fd(q, h=1e-5):
    g = np.zeros_like(q)
    for b in range(q.shape[0]):
        for i in range(4):
            for d in range(3):
                qp, qm = q.copy(), q.copy()
                qp[b, i, d] += h
                qm[b, i, d] -= h
                g[b, i, d] = (ol.egrad_oh3(qp)[0][b] - ol.egrad_oh3(qm)[0][b]) / (2 * h)
    return g


def _max_oh(q):
    return np.linalg.norm(q[:, 1:] - q[:, :1], axis=2).max(axis=1)


def test_oh3_gradient_is_derivative_of_energy():
    # below 37.6 bohr O-H every VH2O switching factor is nonzero
    q = np.concatenate([_cloud(6, 0.2, 2), _special()])
    q = q[_max_oh(q) < 37.0]
    _, g = ol.egrad_oh3(q)
    np.testing.assert_allclose(_fd(q), g, atol=2e-8)


def test_oh3_quirk_is_exercised():
    # where an H2O switching factor is zero the reference keeps the previous
    # routine's derivative; the analytic gradient then departs from the
    # energy's wherever that derivative (of a stretched H2) is not small
    q = _special()
    q = q[_max_oh(q) > 38.0]
    _, g = ol.egrad_oh3(q)
    assert (np.abs(_fd(q) - g).max(axis=(1, 2)) > 1e-3).sum() >= 3


def test_oh3_exchange_symmetry_and_rigid_motion():
    q = _cloud(100, 0.3, 3)
    v, g = ol.egrad_oh3(q)
    vs, gs = ol.egrad_oh3(q[:, [0, 1, 3, 2]])
    np.testing.assert_allclose(vs, v, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(gs, g[:, [0, 1, 3, 2]], rtol=0, atol=1e-13)
    rng = np.random.default_rng(4)
    a, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    vr, gr = ol.egrad_oh3(q @ a.T + rng.normal(size=3))
    np.testing.assert_allclose(vr, v, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(gr, g @ a.T, rtol=0, atol=1e-12)
    # translation and rotation invariance: the total force and torque vanish
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-14)
    np.testing.assert_allclose(np.cross(q, g).sum(axis=1), 0.0, atol=1e-12)


@pytest.mark.parametrize("r_oh,r_hh", [(1.8633, 1.4286), (1.6, 1.3), (2.5, 2.0), (3.5, 1.1)])
def test_oh3_separated_fragments_closed_form(r_oh, r_hh):
    v, g = ol.egrad_oh3(_separated(r_oh, r_hh, 80.0)[None])

    def morse(k, r):
        u = 1.0 - np.exp(-BETA[k] * (r - RE[k]))
        return DE[k] * u * u, 2.0 * BETA[k] * DE[k] * u * (1.0 - u)

    (v_oh, dv_oh), (v_hh, dv_hh) = morse(0, r_oh), morse(2, r_hh)
    assert abs(v[0] - (v_oh + v_hh)) <= 1e-14
    np.testing.assert_allclose(g[0, 1], [dv_oh, 0, 0], atol=1e-14)
    np.testing.assert_allclose(g[0, 3], [0, dv_hh, 0], atol=1e-14)


def test_oh3_saddle():
    v, g = ol.egrad_oh3(workloads.OH3_TS[None])
    assert abs(v[0] * 627.5094743 - 6.0909) < 1e-3
    assert np.abs(g).max() < 1e-12
