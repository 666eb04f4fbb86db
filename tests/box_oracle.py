"""Ring-polymer dynamics of a periodic box on the CPU -- test infrastructure
(the oracle), used only to check the GPU box engine.

One velocity-Verlet RPMD step as the reference's verlet does it
(src/verlet.f90:216-218 half kick, :377-463 free ring-polymer propagation in
normal-mode space, :772-777 forces, :1031-1034 half kick), with the forces of
every bead from the oracle force field (oracle/ff_box.c).  The normal-mode
transform is the reference's rfft/irfft pair: both directions apply the
symmetric matrix cos(2 pi k n / N)/sqrt(N) (src/rfft.f90:36-62,
src/irfft.f90:36-63).  The numpy propagation is pinned against the C
oracle's oracle_free_rp on a small system (tests/test_oracle_box.py).
"""
import numpy as np


def cosine_matrix(nb):
    k = np.arange(nb)
    return np.cos(2.0 * np.pi * np.outer(k, k) / nb) / np.sqrt(nb)


def orthogonal_matrix(nb):
    """Orthogonal real-DFT normal-mode basis: row k = 0 constant, rows
    1..nb/2 cosines, rows above nb/2 sines, so that modes k and nb - k share
    the frequency 2 sin(k pi / nb) / beta_n."""
    n = np.arange(nb)
    u = np.empty((nb, nb))
    u[0] = 1.0 / np.sqrt(nb)
    for k in range(1, nb):
        if k < nb / 2:
            u[k] = np.sqrt(2.0 / nb) * np.cos(2.0 * np.pi * k * n / nb)
        elif k == nb / 2:
            u[k] = (-1.0) ** n / np.sqrt(nb)
        else:
            u[k] = np.sqrt(2.0 / nb) * np.sin(2.0 * np.pi * k * n / nb)
    return u


def propagator(nb, dt, masses, beta):
    """poly[c, k, atom] of src/verlet.f90:405-432: (p, q) -> (p c0 + q c1, p c2 + q c3)."""
    na = len(masses)
    poly = np.zeros((4, nb, na))
    poly[0, 0] = 1.0
    poly[2, 0] = dt / masses
    poly[3, 0] = 1.0
    if nb > 1:
        beta_n = beta / nb
        for k in range(1, nb // 2 + 1):
            wk = (2.0 / beta_n) * np.sin(k * np.pi / nb)
            wt = wk * dt
            wm = wk * masses
            poly[0, k] = np.cos(wt)
            poly[1, k] = -wm * np.sin(wt)
            poly[2, k] = np.sin(wt) / wm
            poly[3, k] = np.cos(wt)
        for k in range(1, (nb - 1) // 2 + 1):
            poly[:, nb - k] = poly[:, k]
    return poly


def free_rp(q, p, dt, masses, beta, orthogonal=False):
    """q, p: (nbeads, natoms, 3); returns the propagated pair.  orthogonal
    selects the orthogonal real-DFT basis (inverse = transpose) instead of the
    reference's cosine pair."""
    nb = q.shape[0]
    if nb == 1:
        return q + p * (dt / masses)[None, :, None], p.copy()
    fwd = orthogonal_matrix(nb) if orthogonal else cosine_matrix(nb)
    inv = fwd.T
    poly = propagator(nb, dt, masses, beta)
    pm = np.einsum("kn,naj->kaj", fwd, p)
    qm = np.einsum("kn,naj->kaj", fwd, q)
    pn = pm * poly[0][:, :, None] + qm * poly[1][:, :, None]
    qn = pm * poly[2][:, :, None] + qm * poly[3][:, :, None]
    return np.einsum("kn,naj->kaj", inv, qn), np.einsum("kn,naj->kaj", inv, pn)


def nhc(p, masses, vnh, nose_q, beta, dt):
    """Nose-Hoover chain over half a step (src/nhc.f90:35-169), numpy
    restatement for the box: centroid-momenta kinetic energy, four thermostats
    (masses as mdinit sets them), nc = 5 Suzuki-Yoshida substeps; returns the
    scaled momenta and updates vnh in place.  Pinned against the C oracle's
    oracle_nhc on a small system (tests/test_box_oracle.py)."""
    nb, na = p.shape[0], p.shape[1]
    nfree = 3 * na
    kelvin = 1.0 / (beta * 0.316679e-5)
    qterm = 0.316679e-5 * kelvin * nose_q * nose_q
    qnh = [nfree * qterm, qterm, qterm, qterm]
    ekt = float(np.float32(np.float32(1.380649e-23) / np.float32(4.3597447e-18))) * kelvin
    w0 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
    ws = (w0, 1.0 - 2.0 * w0, w0)
    dtc = dt / 5.0
    eksum = float(np.sum(p * p / (2.0 * masses[None, :, None]))) / nb / nb
    scale = 1.0
    g = [0.0] * 4
    for _ in range(5):
        for wj in ws:
            dts = wj * dtc
            dt2, dt4, dt8 = 0.5 * dts, 0.25 * dts, 0.125 * dts
            g[3] = (qnh[2] * vnh[2] * vnh[2] - ekt) / qnh[3]
            vnh[3] = vnh[3] + g[3] * dt4
            for j in (2, 1):
                g[j] = (qnh[j - 1] * vnh[j - 1] * vnh[j - 1] - ekt) / qnh[j]
                e = np.exp(-vnh[j + 1] * dt8)
                vnh[j] = e * (vnh[j] * e + g[j] * dt4)
            g[0] = (2.0 * eksum - nfree * ekt) / qnh[0]
            e = np.exp(-vnh[1] * dt8)
            vnh[0] = e * (vnh[0] * e + g[0] * dt4)
            e = np.exp(-vnh[0] * dt2)
            scale *= e
            eksum *= e * e
            g[0] = (2.0 * eksum - nfree * ekt) / qnh[0]
            e = np.exp(-vnh[1] * dt8)
            vnh[0] = e * (vnh[0] * e + g[0] * dt4)
            for j in (1, 2):
                g[j] = (qnh[j - 1] * vnh[j - 1] * vnh[j - 1] - ekt) / qnh[j]
                e = np.exp(-vnh[j + 1] * dt8)
                vnh[j] = e * (vnh[j] * e + g[j] * dt4)
            g[3] = (qnh[2] * vnh[2] * vnh[2] - ekt) / qnh[3]
            vnh[3] = vnh[3] + g[3] * dt4
    return scale * p


def wrap(q, box):
    """The periodic wrap of src/verlet.f90:540-640 (orthorhombic box): bead
    by bead, an atom whose coordinate lies below 0 or above the box length is
    moved by a box length with all its beads, until inside."""
    nb, na, _ = q.shape
    for k in range(nb):
        for j in range(na):
            for c in range(3):
                L = box[c]
                while q[k, j, c] < 0.0:
                    q[:, j, c] += L
                while q[k, j, c] > L:
                    q[:, j, c] -= L


def transrot(q, p, masses):
    """src/transrot.f90:36-237 over a whole box, numpy restatement of the C
    oracle's oracle_transrot (which holds at most 16 atoms; pinned against it
    in tests/test_box_oracle.py): centre-of-mass velocity and centre with the
    reference's total mass nbeads^2 * M, angular momentum about that centre,
    inertia tensor, and the removal of both from every bead's velocity."""
    nb = q.shape[0]
    m = np.asarray(masses, dtype=np.float64)[None, :, None]
    vel = p / m
    totmass = float(np.sum(np.repeat(np.asarray(masses, dtype=np.float64), nb))) * nb
    vtot = np.sum(vel * m, axis=(0, 1)) / totmass
    ctr = np.sum(q * m, axis=(0, 1)) / totmass
    mang = np.sum(np.cross(q, vel) * m, axis=(0, 1)) - np.cross(ctr, vtot) * totmass
    d = q - ctr
    w = np.broadcast_to(m, q.shape)[..., 0]
    xx, yy, zz = (np.sum(d[..., a] * d[..., a] * w) for a in range(3))
    xy, xz, yz = np.sum(d[..., 0] * d[..., 1] * w), np.sum(d[..., 0] * d[..., 2] * w), np.sum(d[..., 1] * d[..., 2] * w)
    ten = np.array([[yy + zz, -xy, -xz], [-xy, xx + zz, -yz], [-xz, -yz, xx + yy]])
    if q.shape[1] <= 2:
        ten = ten + np.eye(3) * 0.000001
    vang = np.linalg.solve(ten, mang)
    v = vel - vtot
    v = v - np.cross(vang, d)
    return v * m


class QmdffFF:
    """The oracle's QMDFF in the interface BoxRPMD expects."""

    def __init__(self, qo):
        self.qo = qo

    def eval(self, theta, x):
        e, g, _ = self.qo.eval(x)
        return e, g, None, None


class BoxRPMD:
    """The oracle's ring polymer of one box; arrays (nbeads, natoms, 3), bohr
    and atomic units.  nose_q > 0 adds the Nose-Hoover chain before the first
    and after the second half kick (src/verlet.f90:185-206, :1086-1106)."""

    def __init__(self, ffo, theta, q, p, dt, masses, beta, orthogonal=False, nose_q=0.0, wrap_box=None,
                 remove_com=False):
        self.orthogonal, self.remove_com = orthogonal, remove_com
        self.wrap_box = None if wrap_box is None else np.asarray(wrap_box, dtype=np.float64)
        self.nose_q, self.vnh = nose_q, [0.0, 0.0, 0.0, 0.0]
        self.ffo, self.theta = ffo, np.asarray(theta, dtype=np.float64)
        self.q, self.p = np.array(q, dtype=np.float64), np.array(p, dtype=np.float64)
        self.dt, self.m, self.beta = dt, np.asarray(masses, dtype=np.float64), beta
        self.v, self.g = self.forces(self.q)

    def forces(self, q):
        v = np.empty(q.shape[0])
        g = np.empty_like(q)
        for b in range(q.shape[0]):
            e, gb, _, _ = self.ffo.eval(self.theta, q[b])
            v[b], g[b] = e, gb
        return v, g

    def step(self):
        hdt = 0.5 * self.dt
        if self.nose_q > 0:
            self.p = nhc(self.p, self.m, self.vnh, self.nose_q, self.beta, self.dt)
        self.p = self.p - hdt * self.g
        self.q, self.p = free_rp(self.q, self.p, self.dt, self.m, self.beta, self.orthogonal)
        if self.wrap_box is not None:
            wrap(self.q, self.wrap_box)
        self.v, self.g = self.forces(self.q)
        self.p = self.p - hdt * self.g
        if self.nose_q > 0:
            self.p = nhc(self.p, self.m, self.vnh, self.nose_q, self.beta, self.dt)
        if self.remove_com:
            self.p = transrot(self.q, self.p, self.m)

    def energies(self):
        """Per-bead potential, kinetic energy and ring-polymer spring energy
        0.5 m omega_n^2 sum_k |q_k - q_k+1|^2 with omega_n = nbeads / beta, the
        spring whose normal modes have the frequencies of src/verlet.f90:413-418."""
        nb = self.q.shape[0]
        ek = 0.5 * np.sum(self.p ** 2 / self.m[None, :, None])
        wn = nb / self.beta
        d = self.q - np.roll(self.q, -1, axis=0)
        es = 0.5 * wn * wn * np.sum(self.m[None, :, None] * d * d) if nb > 1 else 0.0
        return self.v.copy(), ek, es
