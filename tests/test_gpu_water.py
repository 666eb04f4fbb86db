"""GPU parity of the reference's flexible water model (cam_water_create):
energies and gradients against the oracle (itself pinned to the compiled
reference) for a cluster and for periodic boxes with plain and Zahn Coulomb,
mixed precision within 1e-5, and ring-polymer dynamics of a water cluster
against the oracle's ring polymer."""
import numpy as np
import pytest

import caracal_amd as cam
import geometries as geo
import oracle_lib as ol
from caracal_amd import ff as cff

pytestmark = pytest.mark.gpu
BOHR = 0.52917721092


class _WaterOracle:
    """Oracle adapter with the force-field oracle's eval signature."""

    def __init__(self, **kw):
        self.kw = kw

    def eval(self, theta, x, v=None, want_grad=True):
        e, g = ol.water_eval(x, **self.kw)
        return e, g, None, None


def _opts(kind, box):
    if kind == "cluster":
        return dict(), dict()
    rc = 8.0 / BOHR
    zahn = kind == "zahn"
    o = dict(periodic=True, box=box, zahn=zahn, cut=rc)
    if zahn:
        from math import erfc, exp, pi, sqrt
        za = 0.2 * BOHR
        o.update(zahn_a=za, zahn_par=erfc(za * rc) / rc ** 2 + 2 * za / sqrt(pi) * exp(-(za * rc) ** 2) / rc)
    return o, dict(periodic=True, box=box, zahn=zahn, rcut=rc)


@pytest.mark.parametrize("kind", ["cluster", "plain", "zahn"])
def test_water_matches_oracle(kind):
    import torch
    nw = 27 if kind == "cluster" else 343
    box = 12.0 / BOHR if kind == "cluster" else 22.0 / BOHR
    ko, kf = _opts(kind, box)
    nb = 3
    xs = np.stack([geo.water_box(nw, box, 20 + b) for b in range(nb)])
    ff = cff.ForceField.water(nw, **kf)
    x = torch.tensor(xs, device="cuda")
    e, g, _, _ = ff.forces(None, x)
    e, g = e.cpu().numpy(), g.cpu().numpy()
    for b in range(nb):
        er, gr = ol.water_eval(xs[b], **ko)
        assert abs(e[b] - er) <= 1e-11 * abs(er)
        np.testing.assert_allclose(g[b], gr, rtol=0, atol=1e-11 * np.abs(gr).max())
    em = ff.energy(None, x, precision=cam.CAM_PREC_MIXED).cpu().numpy()
    _, gm, _, _ = ff.forces(None, x, precision=cam.CAM_PREC_MIXED)
    ferr = np.abs(gm.cpu().numpy() - g).max() / np.abs(g).max()
    print(f"water {kind}: mixed-precision force error relative to max |F| {ferr:.2e}")
    assert ferr <= 1e-5
    assert np.max(np.abs(em - e) / np.abs(e)) < 1e-5
    ff.close()


def test_water_cluster_ring_polymer_matches_oracle():
    """27-molecule cluster, 8 beads, FP64: 30 steps of the box engine on the
    water model against the oracle's ring polymer."""
    import box_oracle as bo
    nw, nb = 27, 8
    x = geo.water_box(nw, 12.0 / BOHR, 5, jitter=0.02)
    m = np.tile(cam.mass_au([15.99491461956, 1.00782503207, 1.00782503207]), nw)
    beta = float(cam.beta_of(300.0))
    rng = np.random.default_rng(3)
    q = x[None] + rng.normal(0, 0.02, (nb,) + x.shape)
    p = rng.normal(size=q.shape) * np.sqrt(m[None, :, None] * nb / beta)
    ff = cff.ForceField.water(nw)
    dt = 0.1 * cam.FS_TO_AU
    eng = cff.BoxEngine(ff, nb, dt, m, beta, None, precision=cam.CAM_PREC_FP64, nm_mode=cam.CAM_NM_REFERENCE)
    eng.set_state(q, p)
    t = bo.BoxRPMD(_WaterOracle(), np.zeros(10), q, p, dt, m, beta)
    for block in range(3):
        eng.run(10)
        for _ in range(10):
            t.step()
        qg, _ = eng.get_state()
        err = np.abs(qg - t.q).max() / np.abs(t.q).max()
        print(f"water cluster step {10 * (block + 1)}: max rel position error {err:.2e}")
        assert err < 1e-10
    eng.close()
    ff.close()
