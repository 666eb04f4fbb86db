"""Large batches: a trajectory's result does not depend on the batch it sits
in (its block, its place in the block, the size of the grid).  Twenty
trajectories run alone and embedded at the start, middle and end of a batch
of 100,003 (CH4H, 8 beads: 1.2e5 x 8 x 18 doubles = 115 MB per state array),
through the folded and the unfolded step kernels; bit-for-bit equality."""
import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nmk", [cam.CAM_NMK_FOLDED, cam.CAM_NMK_VALU])
def test_results_do_not_depend_on_batch_position(nmk):
    nb, big, few = 8, 100_003, 20
    w = workloads.ch4h_children(ntraj=few, nbeads=nb, seed=11)
    rng = np.random.default_rng(11)
    p = rng.normal(size=w["q"].shape) * np.sqrt(w["masses"][None, None, :, None] * nb / w["beta"][:, None, None, None])
    kw = dict(precision=cam.CAM_PREC_FP64, ensemble=cam.CAM_ENS_NVE, steps_per_launch=10, nm_kernel=nmk)

    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, few, w["dt"], w["masses"], **kw)
    e.set_state(w["q"], p, w["beta"])
    e.run(20)
    q_alone, p_alone = e.get_state()
    e.close()

    filler = workloads.ch4h_children(ntraj=big, nbeads=nb, seed=12)
    qb, pb, bb = filler["q"], np.zeros_like(filler["q"]), filler["beta"].copy()
    slots = [0, big // 2 - 7, big - few]      # first block, a block boundary region, the last (partial) block
    for s0 in slots:
        qb[s0:s0 + few], pb[s0:s0 + few], bb[s0:s0 + few] = w["q"], p, w["beta"]
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, big, w["dt"], w["masses"], **kw)
    e.set_state(qb, pb, bb)
    e.run(20)
    qo, po = e.get_state()
    e.close()
    for s0 in slots:
        assert np.array_equal(qo[s0:s0 + few], q_alone), s0
        assert np.array_equal(po[s0:s0 + few], p_alone), s0
    assert np.isfinite(qo).all()
