"""Trajectory output of the reference's MD step (src/verlet.f90:783-957), written
on the host from engine snapshots.  The engine keeps the state in HBM; a caller
that wants the reference's files runs the engine iwrite steps at a time and
hands each snapshot (get_state / get_forces) to these writers, which matches
the reference's mod(istep, iwrite) == 0 frames.

Files and their layout, as the reference writes them:

trajectory.xyz      natoms*nbeads / blank line (box lengths in angstrom under
                    NPT) / one line per bead and atom: name x y z (angstrom,
                    bead-major) (src/verlet.f90:789-799, opened by
                    src/dynamic.f90:1480)
md_gradients.dat    natoms*nbeads / blank / atom bead dV/dx dV/dy dV/dz
                    (hartree/bohr) (src/verlet.f90:939-944)
md_velocities.dat   natoms*nbeads / blank / atom bead vx vy vz (p/m, atomic
                    units) (src/verlet.f90:953-958)
traj_gen.xyz,       natoms / "Generation step No. istep" / name x y z of bead 1
traj_cross.xyz      (angstrom); with traj_gen_xi.dat / traj_cross_xi.dat: xi_real
                    xi_ideal per frame (src/verlet.f90:897-913,
                    src/calc_rate.f90:820-824 and 2211-2215)
MACE frames         natoms / Properties=species:S:1:pos:R:3:REF_forces:R:3
                    REF_energy=E pbc="F F F" / name, positions (angstrom, f14.8)
                    and forces (eV/angstrom, f14.8) of bead 1; E in eV
                    (src/verlet.f90:883-893)

CONTCAR,            for structures read from a VASP POSCAR (one bead only,
XDATCAR[_gen|_cross] src/verlet.f90:806-880): CONTCAR rewritten at every frame
                    (title with the step, scale, cell vectors, element names
                    and counts, "Selective dynamics" if the POSCAR had it,
                    "Direct", fractional positions with T/F move flags);
                    XDATCAR the same header once, then "Direct step istep"
                    and the fractional positions per frame (XDATCAR_gen /
                    XDATCAR_cross under calc_rate, src/calc_rate.f90:820-824,
                    :2211-2215).  Fractional positions are inverse(a b c)
                    times the Cartesian positions in angstrom, the inverse by
                    the cofactor formula of src/matinv3.f90 (the cell matrix
                    without the POSCAR scale factor, as the reference has
                    it).  With MACE output the frame line carries the energy
                    in eV (f21.9) and every atom its forces in eV/angstrom.

The reference writes free-form records with list-directed output, whose
spacing depends on its compiler; these writers keep the records and their
fields and print every list-directed real with 17 significant digits, so a
frame reads back to the last bit.  The fixed formats (the MACE lines, the
VASP title and step lines) are reproduced exactly.
"""
import os

import numpy as np

BOHR = 0.52917721092       # angstrom per bohr (src/general.f90:256)
EVOLT = 27.21138503        # eV per hartree (src/general.f90:258)


def _r(x):
    return f"{float(x):25.17E}"


def _i(n):
    return f"{int(n):12d}"


def cell_inverse(a_vec, b_vec, c_vec):
    """Inverse of the matrix with columns a, b, c by the cofactor formula of
    src/matinv3.f90:31-45, term by term."""
    A = np.column_stack([np.asarray(v, dtype=np.float64) for v in (a_vec, b_vec, c_vec)])
    detinv = 1.0 / (A[0, 0] * A[1, 1] * A[2, 2] - A[0, 0] * A[1, 2] * A[2, 1]
                    - A[0, 1] * A[1, 0] * A[2, 2] + A[0, 1] * A[1, 2] * A[2, 0]
                    + A[0, 2] * A[1, 0] * A[2, 1] - A[0, 2] * A[1, 1] * A[2, 0])
    B = np.empty((3, 3))
    B[0, 0] = +detinv * (A[1, 1] * A[2, 2] - A[1, 2] * A[2, 1])
    B[1, 0] = -detinv * (A[1, 0] * A[2, 2] - A[1, 2] * A[2, 0])
    B[2, 0] = +detinv * (A[1, 0] * A[2, 1] - A[1, 1] * A[2, 0])
    B[0, 1] = -detinv * (A[0, 1] * A[2, 2] - A[0, 2] * A[2, 1])
    B[1, 1] = +detinv * (A[0, 0] * A[2, 2] - A[0, 2] * A[2, 0])
    B[2, 1] = -detinv * (A[0, 0] * A[2, 1] - A[0, 1] * A[2, 0])
    B[0, 2] = +detinv * (A[0, 1] * A[1, 2] - A[0, 2] * A[1, 1])
    B[1, 2] = -detinv * (A[0, 0] * A[1, 2] - A[0, 2] * A[1, 0])
    B[2, 2] = +detinv * (A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0])
    return B


class VaspCell:
    """What the reference keeps from a POSCAR for its output: scale factor,
    cell vectors a, b, c (angstrom, before scaling), element names and
    counts, selective dynamics and the per-atom move flags (True: T T T)."""

    def __init__(self, scale, a_vec, b_vec, c_vec, names, numbers, selective=False, move=None):
        self.scale = float(scale)
        self.vecs = [np.asarray(v, dtype=np.float64) for v in (a_vec, b_vec, c_vec)]
        self.names, self.numbers = list(names), [int(n) for n in numbers]
        self.selective = bool(selective)
        self.move = None if move is None else [bool(m) for m in move]
        self.inv = cell_inverse(*self.vecs)

    def header(self):
        out = [_r(self.scale)] + ["".join(_r(x) for x in v) for v in self.vecs]
        out.append("".join(" " + n.strip() for n in self.names))
        out.append("".join(_i(n) for n in self.numbers))
        return "\n".join(out) + "\n"

    def fractional(self, x_angstrom):
        """matmul(coord_inv, x) per atom, summed in index order."""
        B = self.inv
        x = np.asarray(x_angstrom, dtype=np.float64)
        return np.stack([(B[i, 0] * x[:, 0] + B[i, 1] * x[:, 1]) + B[i, 2] * x[:, 2] for i in range(3)], axis=1)


def _frames(a, natoms):
    """(nbeads, natoms, 3) view of one trajectory's beads (a single (natoms, 3) frame is one bead)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2:
        a = a[None]
    if a.shape[-1] != 3 or a.shape[-2] != natoms:
        raise ValueError(f"expected (nbeads, {natoms}, 3), got {a.shape}")
    return a


class TrajectoryWriter:
    """Writers of one trajectory's files.  names: element symbols of the atoms.
    Each write_* call appends one frame; flush() or close() ends the files."""

    def __init__(self, names, directory=".", *, verbose=False, box=None, vasp=None):
        self.names = list(names)
        self.natoms = len(self.names)
        self.dir = directory
        self.box = None if box is None else np.asarray(box, dtype=np.float64)
        self.vasp = vasp
        self._xdat_first = {}
        self._files = {}
        self._open = lambda name: self._files.setdefault(name, open(os.path.join(self.dir, name), "w"))
        self.verbose = verbose

    # -- dynamic.x ---------------------------------------------------------------
    def write_positions(self, q):
        """One trajectory.xyz frame; q in bohr, (nbeads, natoms, 3)."""
        q = _frames(q, self.natoms)
        f = self._open("trajectory.xyz")
        f.write(f"{self.natoms * q.shape[0]:12d}\n")
        f.write("" if self.box is None else " ".join(_r(b * BOHR) for b in self.box))
        f.write("\n")
        for k in range(q.shape[0]):
            for i in range(self.natoms):
                f.write(f" {self.names[i]:<2s}" + "".join(_r(x * BOHR) for x in q[k, i]) + "\n")

    def _indexed(self, name, a):
        a = _frames(a, self.natoms)
        f = self._open(name)
        f.write(f"{self.natoms * a.shape[0]:12d}\n\n")
        for k in range(a.shape[0]):
            for i in range(self.natoms):
                f.write(f"{i + 1:12d}{k + 1:12d}" + "".join(_r(x) for x in a[k, i]) + "\n")

    def write_gradients(self, g):
        """One md_gradients.dat frame (verbose runs); g in hartree/bohr."""
        self._indexed("md_gradients.dat", g)

    def write_velocities(self, p, masses):
        """One md_velocities.dat frame (verbose runs): p / m in atomic units."""
        p = _frames(p, self.natoms)
        self._indexed("md_velocities.dat", p / np.asarray(masses, dtype=np.float64)[None, :, None])

    def write_md_frame(self, q, g=None, p=None, masses=None):
        """Everything the reference writes at a frame of dynamic.x."""
        self.write_positions(q)
        if self.verbose:
            self.write_gradients(g)
            self.write_velocities(p, masses)

    # -- calc_rate.x (print_gen / print_cross) -------------------------------------
    def write_rate_frame(self, q, istep, xi_real, xi_ideal, kind="gen"):
        """One traj_{gen,cross}.xyz frame (bead 1) and its xi line."""
        if kind not in ("gen", "cross"):
            raise ValueError("kind is 'gen' or 'cross'")
        q = _frames(q, self.natoms)
        f = self._open(f"traj_{kind}.xyz")
        f.write(f"{self.natoms:12d}\n Generation step No.{istep:12d}\n")
        for i in range(self.natoms):
            f.write(f" {self.names[i]:<2s}" + "".join(_r(x * BOHR) for x in q[0, i]) + "\n")
        self._open(f"traj_{kind}_xi.dat").write(_r(xi_real) + _r(xi_ideal) + "\n")

    def write_mace_frame(self, q, g, epot, xi=None, kind="gen"):
        """One MACE training-set frame of bead 1: energy in eV, forces -g in
        eV/angstrom (train_format MACE).  xi = (xi_real, xi_ideal) also
        writes the xi line, as under calc_rate."""
        q, g = _frames(q, self.natoms), _frames(g, self.natoms)
        f = self._open(f"traj_{kind}.xyz")
        f.write(f"{self.natoms:12d}\n")
        f.write(f"Properties=species:S:1:pos:R:3:REF_forces:R:3 REF_energy={epot * EVOLT:20.10f} pbc=\"F F F\"\n")
        for i in range(self.natoms):
            pos = "".join(f"{x * BOHR:14.8f}" for x in q[0, i])
            frc = "".join(f"{-x / BOHR * EVOLT:14.8f}" for x in g[0, i])
            f.write(f"{self.names[i]}  {pos} {frc}\n")
        if xi is not None:
            self._open(f"traj_{kind}_xi.dat").write(_r(xi[0]) + _r(xi[1]) + "\n")

    # -- VASP input (coord_vasp) --------------------------------------------------
    def write_vasp_frame(self, q, istep, kind="md", xi=None, g=None, epot=None):
        """One CONTCAR / XDATCAR frame of a one-bead run whose structure came
        from a POSCAR (src/verlet.f90:806-880).  kind "md" (dynamic.x:
        XDATCAR), "gen" or "cross" (calc_rate: XDATCAR_gen, XDATCAR_cross).
        xi = (xi_real, xi_ideal) also writes the xi line (calc_rate).  With
        g and epot the frame is the MACE variant (energy on the frame line,
        forces in eV/angstrom after the positions)."""
        v = self.vasp
        if v is None:
            raise ValueError("the writer has no VaspCell")
        q = _frames(q, self.natoms)
        if q.shape[0] != 1:
            raise ValueError("the VASP frames are written for one bead only")
        if kind not in ("md", "gen", "cross"):
            raise ValueError("kind is 'md', 'gen' or 'cross'")
        mace = g is not None
        frac = v.fractional(q[0] * BOHR)
        forces = None if not mace else -_frames(g, self.natoms)[0] / BOHR * EVOLT
        prog = "dynamic.x" if kind == "md" else "calc_rate.x"
        with open(os.path.join(self.dir, "CONTCAR"), "w") as c:
            c.write(f"CONTCAR (step {int(istep):9d}), written by Caracal (dynamic.x)\n")
            c.write(v.header())
            if v.selective:
                c.write(" Selective dynamics\n")
            c.write(" Direct\n")
            xname = {"md": "XDATCAR", "gen": "XDATCAR_gen", "cross": "XDATCAR_cross"}[kind]
            x = self._open(xname)
            if self._xdat_first.setdefault(xname, True):
                x.write(f"step {int(istep):9d}, written by Caracal ({prog})\n")
                x.write(v.header())
                self._xdat_first[xname] = False
            if xi is not None:
                self._open(f"traj_{kind}_xi.dat").write(_r(xi[0]) + _r(xi[1]) + "\n")
            if mace:
                x.write(f"Direct step {int(istep):9d} energy: {epot * EVOLT:21.9f}\n")
                for i in range(self.natoms):
                    row = "".join(_r(t) for t in frac[i]) + "".join(_r(t) for t in forces[i]) + "\n"
                    c.write(row)
                    x.write(row)
            else:
                x.write(f"Direct step {int(istep):9d}\n")
                for i in range(self.natoms):
                    row = "".join(_r(t) for t in frac[i])
                    flag = ""
                    if v.selective:
                        flag = "   T   T   T " if (v.move is None or v.move[i]) else "   F   F   F "
                    c.write(row + flag + "\n")
                    x.write(row + "\n")

    def flush(self):
        for f in self._files.values():
            f.flush()

    def close(self):
        for f in self._files.values():
            f.close()
        self._files.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_xyz_frames(path):
    """Frames of an xyz file written above: list of (comment, names, (n, 3) array)."""
    out = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i])
        comment = lines[i + 1]
        rows = [lines[i + 2 + j].split() for j in range(n)]
        out.append((comment, [r[0] for r in rows], np.array([[float(x) for x in r[1:4]] for r in rows])))
        i += 2 + n
    return out


def read_indexed_frames(path):
    """Frames of md_gradients.dat / md_velocities.dat: list of (n, 5) arrays
    (atom, bead, x, y, z)."""
    out = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i])
        out.append(np.array([[float(x) for x in lines[i + 2 + j].split()] for j in range(n)]))
        i += 2 + n
    return out
