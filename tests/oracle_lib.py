"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the compiled
reference (oracle/_ref/libcaracal_ref.so).  Test infrastructure only: the
product package never imports this module."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcaracal_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _oracle = ctypes.CDLL(ORACLE_SO)
    return _oracle


def ref():
    """The compiled reference, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
        _ref.initialize_ch4h_()
        _ref.initialize_oh3_()
        _ref.initialize_brh2_()
    return _ref


def _egrad(fn, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    nb, na, _ = q.shape                      # numpy (nbeads, natoms, 3) == Fortran q(3,natoms,nbeads)
    v = np.zeros(nb)
    g = np.zeros_like(q)
    fn(_ptr(q), na, nb, _ptr(v), _ptr(g))
    return v, g


def egrad_h3(q):
    return _egrad(oracle().oracle_egrad_h3, q)


def egrad_ch4h(q):
    return _egrad(oracle().oracle_egrad_ch4h, q)


def egrad_oh3(q):
    return _egrad(oracle().oracle_egrad_oh3, q)


def egrad_brh2(q):
    return _egrad(oracle().oracle_egrad_brh2, q)


def egrad_o3(q):
    return _egrad(oracle().oracle_egrad_o3, q)


def _ref_egrad(fn, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    nb, na, _ = q.shape
    v = np.zeros(nb)
    g = np.zeros_like(q)
    n_at, n_b, info = ctypes.c_int(na), ctypes.c_int(nb), ctypes.c_int(0)
    fn(_ptr(q), ctypes.byref(n_at), ctypes.byref(n_b), _ptr(v), _ptr(g), ctypes.byref(info))
    return v, g


def ref_egrad_h3(q):
    return _ref_egrad(ref().egrad_h3_, q)


def ref_egrad_ch4h(q):
    return _ref_egrad(ref().egrad_ch4h_, q)


def ref_egrad_oh3(q):
    """One bead per call: the reference's egrad_oh3 clears the whole gradient
    array inside its bead loop, and gradient.f90 calls it one bead at a time."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = [_ref_egrad(ref().egrad_oh3_, q[b:b + 1]) for b in range(q.shape[0])]
    return np.concatenate([o[0] for o in out]), np.concatenate([o[1] for o in out])


def ref_egrad_brh2(q):
    return _ref_egrad(ref().egrad_brh2_, q)


def ref_egrad_o3(q):
    return _ref_egrad(ref().egrad_o3_, q)


# ---- RPMD step (oracle/rpmd.c) ----
class OracleSys(ctypes.Structure):
    _fields_ = [("surface", ctypes.c_int), ("natoms", ctypes.c_int), ("nbeads", ctypes.c_int),
                ("dt", ctypes.c_double), ("mass", ctypes.c_double * 16)]


class OracleRx(ctypes.Structure):
    _fields_ = [("n_form", ctypes.c_int), ("n_break", ctypes.c_int),
                ("bond_form", (ctypes.c_int * 2) * 4), ("bond_break", (ctypes.c_int * 2) * 4),
                ("form_ref", ctypes.c_double * 4), ("break_ref", ctypes.c_double * 4),
                ("n_reac", ctypes.c_int), ("reac_natoms", ctypes.c_int * 4),
                ("reac_atoms", (ctypes.c_int * 16) * 4),
                ("r_inf", ctypes.c_double), ("xi_ideal", ctypes.c_double)]


def make_sys(surface, nbeads, dt, masses):
    s = OracleSys()
    s.surface = surface
    s.natoms = len(masses)
    s.nbeads = nbeads
    s.dt = dt
    for i, m in enumerate(masses):
        s.mass[i] = m
    return s


def make_rx(reaction):
    """From a caracal_amd.Reaction (same field meaning)."""
    r = OracleRx()
    r.n_form, r.n_break = len(reaction.bond_form), len(reaction.bond_break)
    for i, (a, b) in enumerate(reaction.bond_form):
        r.bond_form[i][0], r.bond_form[i][1] = a, b
        r.form_ref[i] = reaction.form_ref[i]
    for i, (a, b) in enumerate(reaction.bond_break):
        r.bond_break[i][0], r.bond_break[i][1] = a, b
        r.break_ref[i] = reaction.break_ref[i]
    r.n_reac = len(reaction.reactants)
    for k, atoms in enumerate(reaction.reactants):
        r.reac_natoms[k] = len(atoms)
        for i, a in enumerate(atoms):
            r.reac_atoms[k][i] = a
    r.r_inf, r.xi_ideal = reaction.r_inf, reaction.xi_ideal
    return r


def _setup_rpmd():
    o = oracle()
    o.oracle_verlet.restype = ctypes.c_double
    o.oracle_verlet.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, _dp, _dp,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_int,
                                ctypes.c_uint64, ctypes.c_uint32]
    o.oracle_calc_xi.restype = ctypes.c_double
    o.oracle_calc_xi.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, ctypes.c_int]
    o.oracle_child_weight.restype = ctypes.c_double
    o.oracle_child_weight.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, ctypes.c_double]
    o.oracle_andersen_normal.restype = ctypes.c_double
    o.oracle_andersen_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    o.oracle_andersen.argtypes = [ctypes.POINTER(OracleSys), _dp, ctypes.c_double, ctypes.c_uint64,
                                  ctypes.c_uint32, ctypes.c_uint64]
    o.oracle_constrain_q.restype = ctypes.c_int
    o.oracle_constrain_q.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, _dp, _dp]
    o.oracle_constrain_p.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp]
    o.oracle_parent_begin.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_double]
    o.oracle_calc_xi_hess.restype = ctypes.c_double
    o.oracle_calc_xi_hess.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, _dp, ctypes.c_int]
    o.oracle_umbrella.restype = ctypes.c_double
    o.oracle_umbrella.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, ctypes.c_double,
                                  ctypes.c_double, _dp, _dp, _dp]
    o.oracle_umbrella_integration.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                              _dp]
    o.oracle_pmf_extrema.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip, _ip]
    o.oracle_rate_bimolecular.restype = ctypes.c_double
    o.oracle_rate_bimolecular.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    o.oracle_nhc.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp, _dp, ctypes.c_double, _dp]
    o.oracle_verlet_nhc.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
    o.oracle_verlet_umbrella.restype = ctypes.c_double
    o.oracle_verlet_umbrella.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, _dp, _dp,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_long, ctypes.c_int,
                                         ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int]
    o.oracle_verlet_parent.restype = ctypes.c_int
    o.oracle_verlet_parent.argtypes = [ctypes.POINTER(OracleSys), ctypes.POINTER(OracleRx), _dp, _dp, _dp, _dp, _dp,
                                       ctypes.c_double, ctypes.c_long, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_uint32]
    o.oracle_free_rp.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp, ctypes.c_double]
    o.oracle_rfft.argtypes = [_dp, ctypes.c_int]
    o.oracle_transrot.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp]
    o.oracle_centroid.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp]
    o.oracle_forces.argtypes = [ctypes.POINTER(OracleSys), _dp, _dp, _dp]
    o.oracle_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint32)]
    return o


_rpmd = None


def rpmd():
    global _rpmd
    if _rpmd is None:
        _rpmd = _setup_rpmd()
    return _rpmd


def forces(sys, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.zeros(sys.nbeads)
    g = np.zeros_like(q)
    rpmd().oracle_forces(ctypes.byref(sys), _ptr(q), _ptr(v), _ptr(g))
    return v, g


class OracleTraj:
    """One trajectory stepped by the oracle: q, p (nbeads, natoms, 3)."""

    def __init__(self, sys, q, p, beta, rx=None, ensemble=0, remove_com=False, andersen_step=0, seed=0, traj=0):
        self.sys, self.rx, self.beta = sys, rx if rx is not None else OracleRx(), beta
        self.q = np.array(q, dtype=np.float64, order="C")
        self.p = np.array(p, dtype=np.float64, order="C")
        self.v, self.g = forces(sys, self.q)
        self.ensemble, self.remove_com = ensemble, remove_com
        self.andersen_step, self.seed, self.traj = andersen_step, seed, traj
        self.istep = 0

    def step(self):
        self.istep += 1
        if self.ensemble == 5:
            rpmd().oracle_verlet_nhc(ctypes.byref(self.sys), _ptr(self.q), _ptr(self.p), _ptr(self.g), _ptr(self.v),
                                     self.beta, _ptr(self.vnh), _ptr(self.qnh), _ptr(self.eta))
            if self.remove_com:              # dynamic.x ends every step with transrot
                rpmd().oracle_transrot(ctypes.byref(self.sys), _ptr(self.q), _ptr(self.p))
            return 0.0
        if self.ensemble == 4:
            xi = rpmd().oracle_verlet_umbrella(ctypes.byref(self.sys), ctypes.byref(self.rx), _ptr(self.q),
                                               _ptr(self.p), _ptr(self.g), _ptr(self.v), self.beta, self.k_force,
                                               self.istep, self.andersen_step, self.seed, self.traj,
                                               self.key_base + self.istep, int(self.remove_com))
            self.xi_sum += xi
            self.xi_sq += xi * xi
            return xi
        if self.ensemble == 3:
            return rpmd().oracle_verlet_parent(ctypes.byref(self.sys), ctypes.byref(self.rx), _ptr(self.q),
                                               _ptr(self.p), _ptr(self.g), _ptr(self.v), _ptr(self.dxi), self.beta,
                                               self.istep, self.andersen_step, self.seed, self.traj)
        return rpmd().oracle_verlet(ctypes.byref(self.sys), ctypes.byref(self.rx), _ptr(self.q), _ptr(self.p),
                                    _ptr(self.g), _ptr(self.v), self.beta, self.ensemble, int(self.remove_com),
                                    self.istep, self.andersen_step, self.seed, self.traj)

    def nhc_begin(self, nose_q):
        """Nose-Hoover chain (ensemble 5), as mdinit sets it up
        (src/mdinit.f90:132-146): thermostat velocities 0, masses
        nfree kT Q^2 and kT Q^2 with kT = 0.316679e-5 T; positions 0 (tracked
        for the conserved energy only)."""
        kelvin = 1.0 / (self.beta * 0.316679e-5)
        qterm = 0.316679e-5 * kelvin * nose_q * nose_q
        nfree = 3 * self.q.shape[1]
        self.qnh = np.array([nfree * qterm, qterm, qterm, qterm])
        self.vnh = np.zeros(4)
        self.eta = np.zeros(4)

    def umbrella_begin(self, k_force, bias=True):
        """Umbrella window (ensemble 4; the window centre is self.rx.xi_ideal):
        fresh forces, plus the bias if bias (mdinit's start), and a new
        segment: sums cleared, the Andersen schedule from step 1, draw ids
        continuing."""
        self.k_force = k_force
        self.v, self.g = forces(self.sys, self.q)
        if bias:
            c = np.ascontiguousarray(self.q.mean(axis=0))
            c[:] = 0.0
            rpmd().oracle_centroid(ctypes.byref(self.sys), _ptr(self.q), _ptr(c))
            d = np.zeros_like(c)
            rpmd().oracle_umbrella(ctypes.byref(self.sys), ctypes.byref(self.rx), _ptr(c), self.beta, k_force,
                                   _ptr(d), _ptr(self.g), None)
        self.key_base = getattr(self, "key_base", 0) + self.istep
        self.istep = 0
        self.xi_sum = self.xi_sq = 0.0

    def parent_begin(self, k_force=0.0):
        """Recrossing parent (ensemble 3): mdinit with bias_mode 2 -- fresh
        forces plus the umbrella bias gradient, and the constraint direction."""
        self.v, self.g = forces(self.sys, self.q)
        self.dxi = np.zeros((self.q.shape[1], 3))
        rpmd().oracle_parent_begin(ctypes.byref(self.sys), ctypes.byref(self.rx), _ptr(self.q), _ptr(self.g),
                                   _ptr(self.dxi), self.beta, k_force)


def calc_xi(sys, rx, c, mode=2):
    c = np.ascontiguousarray(c, dtype=np.float64)
    d = np.zeros_like(c)
    xi = rpmd().oracle_calc_xi(ctypes.byref(sys), ctypes.byref(rx), _ptr(c), _ptr(d), mode)
    return xi, d


def calc_xi_hess(sys, rx, c, mode=1):
    c = np.ascontiguousarray(c, dtype=np.float64)
    d = np.zeros_like(c)
    h = np.zeros((c.size, c.size))
    xi = rpmd().oracle_calc_xi_hess(ctypes.byref(sys), ctypes.byref(rx), _ptr(c), _ptr(d), _ptr(h), mode)
    return xi, d, h


def umbrella(sys, rx, c, beta, k_force):
    """umbrella mode 0 at centroid c: (bias energy, xi, dxi, bias gradient of each of sys.nbeads beads)."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    d = np.zeros_like(c)
    g = np.zeros((sys.nbeads,) + c.shape)
    xi = ctypes.c_double()
    v = rpmd().oracle_umbrella(ctypes.byref(sys), ctypes.byref(rx), _ptr(c), beta, k_force, _ptr(d), _ptr(g),
                               ctypes.byref(xi))
    return v, xi.value, d, g


def umbrella_integration(xi_wins, average, variance, k_force, beta, xi_min, xi_max, nbins, umbr_traj=1, umbr_step=1):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (xi_wins, average, variance, k_force)]
    dA, bc, pmf = np.zeros(nbins), np.zeros(nbins - 1), np.zeros(nbins)
    rpmd().oracle_umbrella_integration(len(a[0]), *[_ptr(x) for x in a], beta, xi_min, xi_max, nbins, umbr_traj,
                                       umbr_step, _ptr(dA), _ptr(bc), _ptr(pmf))
    return dA, bc, pmf


def pmf_extrema(pmf, xi_min, xi_max, minloc_zero=True):
    pmf = np.ascontiguousarray(pmf, dtype=np.float64)
    imax, imin = ctypes.c_int(), ctypes.c_int()
    rpmd().oracle_pmf_extrema(_ptr(pmf), len(pmf), xi_min, xi_max, int(minloc_zero), ctypes.byref(imax),
                              ctypes.byref(imin))
    return imax.value, imin.value


def rate_bimolecular(npaths, kappa, r_inf, beta, m1, m2, pmf_max, pmf_min):
    return rpmd().oracle_rate_bimolecular(npaths, kappa, r_inf, beta, m1, m2, pmf_max, pmf_min)


def andersen_normal(seed, traj, draw, coord, bead):
    return rpmd().oracle_andersen_normal(seed, traj, draw, coord, bead)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    rpmd().oracle_philox4x32_10(c, k, o)
    return list(o)


# ---- QMDFF-form box (oracle/ff_box.c) ----
class OracleFF(ctypes.Structure):
    _fields_ = [("natoms", ctypes.c_int), ("nbonds", ctypes.c_int), ("nangles", ctypes.c_int),
                ("type", _ip), ("mol", _ip), ("bonds", _ip), ("angles", _ip),
                ("box", ctypes.c_double), ("rcut", ctypes.c_double), ("zahn_a", ctypes.c_double),
                ("zahn_par", ctypes.c_double),
                ("sr42", (ctypes.c_double * 2) * 2), ("r0bj", (ctypes.c_double * 2) * 2),
                ("zrep", (ctypes.c_double * 2) * 2), ("arep", (ctypes.c_double * 2) * 2)]


class FFOracle:
    def __init__(self, system, rc_angstrom=10.0):
        from caracal_amd import ffbox
        self.keep = [np.ascontiguousarray(system[k], dtype=np.int32) for k in ("type", "mol", "bonds", "angles")]
        f = OracleFF()
        f.natoms = len(system["type"])
        f.nbonds = len(system["bonds"])
        f.nangles = len(system["angles"])
        f.type, f.mol, f.bonds, f.angles = (a.ctypes.data_as(_ip) for a in self.keep)
        f.box = system["box"]
        za, rc, zp = ffbox.zahn_params(rc_angstrom)
        f.rcut, f.zahn_a, f.zahn_par = rc, za, zp
        sr42, r0, zrep, arep = ffbox.pair_tables()
        for a in range(2):
            for b in range(2):
                f.sr42[a][b], f.r0bj[a][b], f.zrep[a][b], f.arep[a][b] = sr42[a, b], r0[a, b], zrep[a, b], arep[a, b]
        self.f = f
        o = oracle()
        o.oracle_ff_eval.restype = ctypes.c_double
        o.oracle_ff_eval.argtypes = [ctypes.POINTER(OracleFF), _dp, _dp, _dp, _dp, _dp, _dp]

    def eval(self, theta, x, v=None, want_grad=True):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x) if want_grad else None
        dth = np.zeros(10)
        dvth = np.zeros(10) if v is not None else None
        vv = np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
        e = oracle().oracle_ff_eval(ctypes.byref(self.f), _ptr(theta), _ptr(x),
                                    _ptr(g) if g is not None else None, _ptr(dth),
                                    _ptr(vv) if vv is not None else None, _ptr(dvth) if dvth is not None else None)
        return e, g, dth, dvth


# ---- water surface (src/egrad_water.f90) ----
def water_params():
    par = np.zeros(11)
    oracle().oracle_water_params(_ptr(par))
    return par


def _water_args(x, periodic, box, zahn, zahn_a, zahn_par, cut):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return x, np.zeros_like(x), (int(periodic), float(box), int(zahn), float(zahn_a), float(zahn_par), float(cut))


def water_eval(x, periodic=False, box=0.0, zahn=False, zahn_a=0.0, zahn_par=0.0, cut=0.0):
    """x: (natoms, 3) bohr, O,H,H order -> (E, g)."""
    x, g, (ip, bx, iz, za, zp, ct) = _water_args(x, periodic, box, zahn, zahn_a, zahn_par, cut)
    f = oracle().oracle_water_eval
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                  ctypes.c_double, ctypes.c_double]
    e = f(len(x) // 3, _ptr(x), _ptr(g), ip, bx, iz, za, zp, ct)
    return e, g


def ref_water_eval(x, periodic=False, box=0.0, zahn=False, zahn_a=0.0, zahn_par=0.0, cut=0.0):
    """The compiled reference's egrad_water through the harness, or None."""
    r = ref()
    if r is None:
        return None
    x, g, (ip, bx, iz, za, zp, ct) = _water_args(x, periodic, box, zahn, zahn_a, zahn_par, cut)
    e = ctypes.c_double()
    f = r.ref_water_eval
    f.restype = None
    f.argtypes = [ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_double,
                  ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    f(len(x), _ptr(x), _ptr(g), ctypes.byref(e), ip, bx, iz, za, zp, ct)
    return e.value, g


# ---- QMDFF (oracle/qmdff.c) ----
class OracleQmdff(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nbond", ctypes.c_int), ("nangl", ctypes.c_int), ("ntors", ctypes.c_int),
                ("nhb", ctypes.c_int), ("nnci", ctypes.c_int), ("nmols", ctypes.c_int),
                ("at", _ip), ("molnum", _dp), ("q", _dp), ("bond", _ip), ("vbond", _dp), ("angl", _ip),
                ("vangl", _dp), ("tors", _ip), ("vtors", _dp), ("hb", _ip), ("vhb", _dp), ("nci", _ip),
                ("c6xy", _dp), ("r094", _dp), ("sr42", _dp), ("zab", _dp), ("r0ab", _dp),
                ("eps1", ctypes.c_double * 6), ("eps2", ctypes.c_double * 6), ("rad", _dp), ("scalehb", _dp),
                ("periodic", ctypes.c_int), ("zahn", ctypes.c_int), ("box", ctypes.c_double * 3),
                ("coul_cut", ctypes.c_double), ("vdw_cut", ctypes.c_double), ("cut_low", ctypes.c_double),
                ("zahn_a", ctypes.c_double), ("zahn_par", ctypes.c_double)]


class QmdffOracle:
    """The oracle's QMDFF over a parameter dict in the layout of
    tests/refqmdff.py (the arrays prepare leaves in the qmdff module)."""

    def __init__(self, p):
        self.keep = {}

        def arr(name, dtype):
            a = np.ascontiguousarray(p[name], dtype=dtype)
            self.keep[name] = a
            return a.ctypes.data_as(_ip if dtype == np.int32 else _dp)

        s = OracleQmdff()
        s.n, s.nmols = int(p["natoms"]), int(p["nmols"])
        s.nbond, s.nangl, s.ntors = len(p["bond"]), len(p["angl"]), len(p["tors"])
        s.nhb, s.nnci = len(p["hb"]), len(p["nci"])
        for name in ("at", "bond", "angl", "tors", "hb", "nci"):
            setattr(s, name, arr(name, np.int32))
        for name in ("molnum", "q", "vbond", "vangl", "vtors", "vhb", "c6xy", "rad", "scalehb"):
            setattr(s, name, arr(name, np.float64))
        for name in ("r094", "sr42", "zab", "r0ab"):
            setattr(s, name, arr(name, np.float64))
        for k in range(6):
            s.eps1[k], s.eps2[k] = p["eps1"][k], p["eps2"][k]
        s.periodic, s.zahn = int(p["periodic"]), int(p["zahn"])
        for k in range(3):
            s.box[k] = p["box"][k]
        s.coul_cut, s.vdw_cut, s.cut_low = p["coul_cut"], p["vdw_cut"], p["cut_low"]
        s.zahn_a, s.zahn_par = p["zahn_a"], p["zahn_par"]
        self.s = s
        o = oracle()
        o.oracle_qmdff_eval.restype = ctypes.c_double
        o.oracle_qmdff_eval.argtypes = [ctypes.POINTER(OracleQmdff), _dp, _dp, _dp]
        self.fn = o.oracle_qmdff_eval

    def eval(self, x):
        """x (natoms, 3) bohr -> (energy, gradient (natoms, 3), parts (3,))."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        parts = np.zeros(3)
        e = self.fn(ctypes.byref(self.s), _ptr(x), _ptr(g), _ptr(parts))
        return e, g, parts
