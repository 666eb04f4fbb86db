"""The Fortran host path: caracal_amd/fortran/rpmd_driver (iso_c_binding
module caracal_amd_mod) advances an H3, OH + H2, BrH2 or O3 NVE ring polymer on the GPU
through the C ABI; its output must match the oracle's verlet."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
from caracal_amd import workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "caracal_amd", "fortran", "rpmd_driver")


SYSTEMS = {1: (workloads.h3_nve, ol.egrad_h3), 3: (workloads.oh3_nve, ol.egrad_oh3),
           4: (workloads.brh2_nve, ol.egrad_brh2), 5: (workloads.o3_nve, ol.egrad_o3)}


@pytest.mark.parametrize("surface", [1, 3, 4, 5], ids=["H3", "OH3", "BrH2", "O3"])
def test_fortran_driver_matches_oracle(tmp_path, surface):
    if not os.path.exists(DRIVER):
        pytest.fail("caracal_amd/fortran/rpmd_driver is not built")
    w = SYSTEMS[surface][0](nbeads=16)
    na = len(w["masses"])
    nsteps = 100
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([surface, 16, 1, 0, 1, nsteps], dtype="<i4").tofile(f)
        np.array([w["dt"], w["beta"]], dtype="<f8").tofile(f)
        np.asarray(w["masses"], dtype="<f8").tofile(f)
        np.ascontiguousarray(w["q"], dtype="<f8").tofile(f)     # (nbeads,na,3) C order == q(3,na,nbeads)
        np.ascontiguousarray(w["p"], dtype="<f8").tofile(f)
    out = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    data = np.fromfile(fout, dtype="<f8")
    n = 16 * 3 * na
    q, p, v = data[:n].reshape(16, na, 3), data[n:2 * n].reshape(16, na, 3), data[2 * n:]
    s = ol.make_sys(surface, 16, w["dt"], w["masses"])
    t = ol.OracleTraj(s, w["q"], w["p"], w["beta"], ensemble=0, remove_com=True)
    for _ in range(nsteps):
        t.step()
    assert np.abs(q - t.q).max() / np.abs(t.q).max() < 1e-10
    v_ref = SYSTEMS[surface][1](w["q"])[0]
    # O3 energies cancel polynomial terms of up to 1e5 kcal/mol: 1e-12 absolute
    np.testing.assert_allclose(v, v_ref, rtol=1e-12, atol=1e-12 if surface == 5 else 1e-15)
