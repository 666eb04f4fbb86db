"""Trajectory output files of the MD step (caracal_amd.trajout, the records of
src/verlet.f90:783-957): frame layout, atom/bead order, units, and values that
read back to the last bit (list-directed fields) or to the fixed formats'
precision (the MACE frame)."""
import numpy as np
import pytest

from caracal_amd import trajout
from caracal_amd.trajout import BOHR, EVOLT, TrajectoryWriter, read_indexed_frames, read_xyz_frames


def _state(nb=3, na=3, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(nb, na, 3)) * 2, rng.normal(size=(nb, na, 3)) * 1e-2, rng.normal(size=(nb, na, 3))


def test_trajectory_xyz_and_verbose_files(tmp_path):
    masses = np.array([1837.0, 1837.0, 3671.0])
    frames = [_state(seed=s) for s in range(3)]
    with TrajectoryWriter(["H", "H", "D"], tmp_path, verbose=True) as w:
        for q, g, p in frames:
            w.write_md_frame(q, g, p, masses)
    xyz = read_xyz_frames(tmp_path / "trajectory.xyz")
    grads = read_indexed_frames(tmp_path / "md_gradients.dat")
    vels = read_indexed_frames(tmp_path / "md_velocities.dat")
    assert len(xyz) == len(grads) == len(vels) == 3
    for (q, g, p), (comment, names, x), gr, vr in zip(frames, xyz, grads, vels):
        assert comment.strip() == "" and names == ["H", "H", "D"] * 3
        assert np.array_equal(x, (q * BOHR).reshape(-1, 3))                 # bead-major, angstrom
        atom_bead = np.array([(i + 1, k + 1) for k in range(3) for i in range(3)])
        assert np.array_equal(gr[:, :2], atom_bead) and np.array_equal(vr[:, :2], atom_bead)
        assert np.array_equal(gr[:, 2:], g.reshape(-1, 3))
        assert np.array_equal(vr[:, 2:], (p / masses[None, :, None]).reshape(-1, 3))


def test_box_line_under_npt(tmp_path):
    with TrajectoryWriter(["O"], tmp_path, box=[10.0, 11.0, 12.0]) as w:
        w.write_positions(np.zeros((2, 1, 3)))
    (comment, _, _), = read_xyz_frames(tmp_path / "trajectory.xyz")
    assert np.allclose([float(t) for t in comment.split()], np.array([10.0, 11.0, 12.0]) * BOHR, rtol=1e-16)


def test_rate_frames_and_xi_lines(tmp_path):
    q, _, _ = _state(nb=4, na=3, seed=5)
    with TrajectoryWriter(["H", "H", "H"], tmp_path) as w:
        for step in (10, 20):
            w.write_rate_frame(q, step, 0.97, 1.0, kind="cross")
    fr = read_xyz_frames(tmp_path / "traj_cross.xyz")
    assert [c.split()[-1] for c, _, _ in fr] == ["10", "20"]
    assert fr[0][0].split()[:3] == ["Generation", "step", "No."]
    assert np.array_equal(fr[0][2], q[0] * BOHR)                            # bead 1 only
    xi = np.loadtxt(tmp_path / "traj_cross_xi.dat")
    assert np.array_equal(xi, [[0.97, 1.0], [0.97, 1.0]])
    with pytest.raises(ValueError):
        TrajectoryWriter(["H"], tmp_path).write_rate_frame(q[:, :1], 1, 0, 0, kind="other")


def test_mace_frame_fixed_format(tmp_path):
    q, g, _ = _state(nb=2, na=2, seed=7)
    with TrajectoryWriter(["C", "H"], tmp_path) as w:
        w.write_mace_frame(q, g, -1.25, xi=(1.01, 1.0))
    lines = (tmp_path / "traj_gen.xyz").read_text().split("\n")
    assert int(lines[0]) == 2
    head = lines[1]
    assert head.startswith("Properties=species:S:1:pos:R:3:REF_forces:R:3 REF_energy=") and head.endswith('pbc="F F F"')
    assert float(head.split("REF_energy=")[1].split()[0]) == pytest.approx(-1.25 * EVOLT, abs=1e-10)
    for i, line in enumerate(lines[2:4]):
        name, rest = line.split("  ", 1)
        assert len(rest) == 3 * 14 + 1 + 3 * 14                            # (a,a,3f14.8,a,3f14.8)
        vals = np.array([float(t) for t in rest.split()])
        np.testing.assert_allclose(vals[:3], q[0, i] * BOHR, atol=5e-9)
        np.testing.assert_allclose(vals[3:], -g[0, i] / BOHR * EVOLT, atol=5e-9)
    assert np.array_equal(np.loadtxt(tmp_path / "traj_gen_xi.dat"), [1.01, 1.0])


def _vasp_cell():
    a, b, c = [7.1, 0.0, 0.0], [1.3, 6.4, 0.0], [-0.7, 0.9, 8.2]
    return trajout.VaspCell(1.0, a, b, c, ["C", "H"], [1, 3], selective=True, move=[True, False, True, True])


def test_cell_inverse_is_the_inverse():
    v = _vasp_cell()
    A = np.column_stack(v.vecs)
    np.testing.assert_allclose(v.inv @ A, np.eye(3), rtol=0, atol=1e-15)
    np.testing.assert_allclose(trajout.cell_inverse(*np.eye(3)), np.eye(3), rtol=0, atol=0)


def test_vasp_contcar_and_xdatcar_frames(tmp_path):
    """CONTCAR rewritten every frame, XDATCAR with one header and a block per
    frame (src/verlet.f90:806-880); fractional positions times the cell give
    back the Cartesian positions in angstrom."""
    v = _vasp_cell()
    rng = np.random.default_rng(3)
    frames = [rng.uniform(0.0, 10.0, (1, 4, 3)) for _ in range(2)]
    with trajout.TrajectoryWriter(["C", "H", "H", "H"], str(tmp_path), vasp=v) as w:
        for step, q in zip((10, 20), frames):
            w.write_vasp_frame(q, step)
    con = (tmp_path / "CONTCAR").read_text().split("\n")
    assert con[0] == "CONTCAR (step        20), written by Caracal (dynamic.x)"
    assert float(con[1]) == 1.0
    np.testing.assert_array_equal([float(x) for x in con[3].split()], v.vecs[1])
    assert con[5].split() == ["C", "H"] and [int(x) for x in con[6].split()] == [1, 3]
    assert con[7].strip() == "Selective dynamics" and con[8].strip() == "Direct"
    rows = [r.split() for r in con[9:13]]
    assert [r[3:] for r in rows] == [["T"] * 3, ["F"] * 3, ["T"] * 3, ["T"] * 3]
    frac = np.array([[float(x) for x in r[:3]] for r in rows])
    cart = frac @ np.stack(v.vecs)
    np.testing.assert_allclose(cart, frames[1][0] * trajout.BOHR, rtol=0, atol=1e-13)
    xd = (tmp_path / "XDATCAR").read_text().split("\n")
    assert xd[0] == "step        10, written by Caracal (dynamic.x)"
    assert xd.count("Direct step        10") == 1 and xd.count("Direct step        20") == 1
    assert sum(line.startswith("step ") for line in xd) == 1
    i20 = xd.index("Direct step        20")
    frac20 = np.array([[float(x) for x in xd[i20 + 1 + k].split()] for k in range(4)])
    np.testing.assert_array_equal(frac20, frac)


def test_vasp_frames_under_calc_rate_and_mace(tmp_path):
    v = _vasp_cell()
    q = np.full((1, 4, 3), 2.0)
    g = np.arange(12, dtype=np.float64).reshape(1, 4, 3) * 1e-3
    with trajout.TrajectoryWriter(["C", "H", "H", "H"], str(tmp_path), vasp=v) as w:
        w.write_vasp_frame(q, 7, kind="gen", xi=(0.5, 0.6), g=g, epot=-0.25)
    xd = (tmp_path / "XDATCAR_gen").read_text().split("\n")
    assert xd[0] == "step         7, written by Caracal (calc_rate.x)"
    assert xd[7] == f"Direct step         7 energy: {-0.25 * trajout.EVOLT:21.9f}"
    row = [float(x) for x in xd[8].split()]
    assert len(row) == 6
    np.testing.assert_allclose(row[3:], -g[0, 0] / trajout.BOHR * trajout.EVOLT, rtol=1e-15)
    assert [float(x) for x in (tmp_path / "traj_gen_xi.dat").read_text().split()] == [0.5, 0.6]
    with pytest.raises(ValueError):
        trajout.TrajectoryWriter(["C"], str(tmp_path), vasp=v).write_vasp_frame(np.zeros((2, 1, 3)), 1)
