"""Trajectory output files of the MD step (caracal_amd.trajout, the records of
src/verlet.f90:783-957): frame layout, atom/bead order, units, and values that
read back to the last bit (list-directed fields) or to the fixed formats'
precision (the MACE frame)."""
import numpy as np
import pytest

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
