"""Host mirror of the QMDFF-form box: the force field (energy kernel and its
hand-written adjoint), a torch.autograd.Function over them, and the box
ring-polymer engine.  Device arrays are torch tensors (float64, cuda)."""
import ctypes

import numpy as np

from . import _abi, ffbox
from ._abi import check, lib

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


class ForceField:
    """Topology and fixed tables of a box (from ffbox.box_system)."""

    def __init__(self, system, rc_angstrom=10.0):
        self._keep = [np.ascontiguousarray(system[k], dtype=np.int32) for k in ("type", "mol", "bonds", "angles")]
        d = _abi.FFDesc()
        d.natoms = len(system["type"])
        d.nbonds = len(system["bonds"])
        d.nangles = len(system["angles"])
        ip = ctypes.POINTER(ctypes.c_int32)
        d.type, d.mol, d.bonds, d.angles = (a.ctypes.data_as(ip) for a in self._keep)
        d.box = system["box"]
        za, rc, zp = ffbox.zahn_params(rc_angstrom)
        d.rcut, d.zahn_a, d.zahn_par = rc, za, zp
        sr42, r0, zrep, arep = ffbox.pair_tables()
        for a in range(2):
            for b in range(2):
                d.sr42[a][b], d.r0bj[a][b], d.zrep[a][b], d.arep[a][b] = sr42[a, b], r0[a, b], zrep[a, b], arep[a, b]
        h = ctypes.c_void_p()
        check("cam_ff_create", lib().cam_ff_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.natoms = d.natoms

    @classmethod
    def water(cls, nwater, periodic=False, box=0.0, zahn=False, rcut=0.0):
        """The reference's flexible water model (atoms O, H, H per molecule);
        box and rcut in bohr.  theta arguments of the methods are ignored."""
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        check("cam_water_create", lib().cam_water_create(nwater, int(periodic), box, int(zahn), rcut, ctypes.byref(h)))
        self._h, self.natoms, self._keep = h, 3 * nwater, []
        return self

    @classmethod
    def qmdff(cls, p):
        """The reference's QMDFF (ff_eg + ff_nonb + ff_hb) from its module
        arrays: a dict as caracal_amd.qmdff.qmdff_params returns (1-based
        atom indices, element-pair tables [z_a - 1][z_b - 1], bohr).  theta
        arguments of the methods are ignored."""
        self = cls.__new__(cls)
        keep = []

        def arr(v, dtype, ctype):
            a = np.ascontiguousarray(v, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctype))

        d = _abi.QmdffDesc()
        d.natoms = int(p["natoms"])
        d.nbond, d.nangl, d.ntors = len(p["bond"]), len(p["angl"]), len(p["tors"])
        d.nhb, d.nnci = len(p["hb"]), len(p["nci"])
        for name in ("at", "bond", "angl", "tors", "hb", "nci"):
            setattr(d, name, arr(p[name], np.int32, ctypes.c_int32))
        for name in ("molnum", "q", "vbond", "vangl", "vtors", "vhb", "c6xy", "rad", "scalehb"):
            setattr(d, name, arr(p[name], np.float64, ctypes.c_double))
        for name in ("r094", "sr42", "zab", "r0ab"):     # to the module's column-major (94, 94)
            setattr(d, name, arr(np.asarray(p[name]).T, np.float64, ctypes.c_double))
        for k in range(6):
            d.eps1[k], d.eps2[k] = p["eps1"][k], p["eps2"][k]
        d.periodic, d.zahn = int(p["periodic"]), int(p["zahn"])
        for k in range(3):
            d.box[k] = p["box"][k]
        d.coul_cut, d.vdw_cut, d.cut_low = p["coul_cut"], p["vdw_cut"], p["cut_low"]
        d.zahn_a, d.zahn_par = p["zahn_a"], p["zahn_par"]
        h = ctypes.c_void_p()
        check("cam_qmdff_create", lib().cam_qmdff_create(ctypes.byref(d), ctypes.byref(h)))
        self._h, self.natoms, self._keep = h, d.natoms, keep
        return self

    def set_pair_mode(self, mode):
        """_abi.CAM_FF_PAIRS_LIST (default: a neighbour list from the bead
        centroids, the same pairs and order as all pairs) or CAM_FF_PAIRS_ALL."""
        check("cam_ff_set_pair_mode", lib().cam_ff_set_pair_mode(self._h, mode))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cam_ff_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the library handle may already be gone
        try:
            self.close()
        except Exception:
            pass

    def energy(self, theta, x, precision=_abi.CAM_PREC_FP64):
        """x: cuda float64 tensor (nbeads, natoms, 3) -> energies (nbeads,)."""
        import torch
        thp = None if theta is None else _p(np.ascontiguousarray(theta, dtype=np.float64))
        x = x.contiguous()
        nb = x.shape[0]
        e = torch.empty(nb, dtype=torch.float64, device=x.device)
        check("cam_ff_energy", lib().cam_ff_energy(self._h, thp, x.data_ptr(), nb, e.data_ptr(), precision,
                                                  torch.cuda.current_stream().cuda_stream))
        return e

    def forces(self, theta, x, v=None, want_params=False, precision=_abi.CAM_PREC_FP64):
        """-> (e (nb,), dE/dx (nb,natoms,3), dE/dtheta (nb,10) or None, d(v.dE/dx)/dtheta (nb,10) or None)."""
        import torch
        thp = None if theta is None else _p(np.ascontiguousarray(theta, dtype=np.float64))
        x = x.contiguous()
        nb = x.shape[0]
        e = torch.empty(nb, dtype=torch.float64, device=x.device)
        g = torch.empty_like(x)
        dth = torch.empty((nb, 10), dtype=torch.float64, device=x.device) if want_params else None
        dvth = torch.empty((nb, 10), dtype=torch.float64, device=x.device) if v is not None else None
        vptr = v.contiguous().data_ptr() if v is not None else None
        check("cam_ff_forces", lib().cam_ff_forces(self._h, thp, x.data_ptr(), nb, e.data_ptr(), g.data_ptr(),
                                                  dth.data_ptr() if dth is not None else None, vptr,
                                                  dvth.data_ptr() if dvth is not None else None, precision,
                                                  torch.cuda.current_stream().cuda_stream))
        return e, g, dth, dvth


def make_autograd(ff):
    """torch.autograd.Function: (x, theta) -> (E per bead, F = -dE/dx).
    Backward returns the parameter gradient, contracting the incoming force
    gradient with d(dE/dx)/dtheta through the adjoint kernel; positions are
    treated as data (no gradient)."""
    import torch

    class FFEnergyForces(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, theta):
            th = theta.detach().cpu().numpy()
            e, g, _, _ = ff.forces(th, x.detach())
            ctx.save_for_backward(x.detach(), theta.detach())
            return e, -g

        @staticmethod
        def backward(ctx, grad_e, grad_f):
            x, theta = ctx.saved_tensors
            th = theta.cpu().numpy()
            need_v = grad_f is not None
            v = (-grad_f).contiguous() if need_v else None      # F = -dE/dx
            _, _, dth, dvth = ff.forces(th, x, v=v, want_params=True)
            gt = torch.zeros(10, dtype=torch.float64, device=x.device)
            if grad_e is not None:
                gt = gt + (grad_e[:, None] * dth).sum(0)
            if need_v:
                gt = gt + dvth.sum(0)
            if ctx.needs_input_grad[0]:
                raise RuntimeError("position gradients through the force field are not provided")
            return None, gt.to(theta.device)

    return FFEnergyForces.apply


class BoxEngine:
    """Ring polymer of one box; state arrays (nbeads, natoms, 3) host."""

    def __init__(self, ff, nbeads, dt, masses, beta, theta, *, nm_mode=_abi.CAM_NM_REFERENCE,
                 precision=_abi.CAM_PREC_MIXED, andersen_step=0, seed=0):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        h = ctypes.c_void_p()
        check("cam_box_create", lib().cam_box_create(ff._h, nbeads, dt, _p(m), beta, nm_mode, precision,
                                                     andersen_step, seed, ctypes.byref(h)))
        self._h, self.ff, self.nbeads, self.natoms = h, ff, nbeads, ff.natoms
        if theta is not None:
            self.set_params(theta)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cam_box_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the library handle may already be gone
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        check("cam_box_set_params", lib().cam_box_set_params(self._h, _p(th)))

    def set_state(self, q, p):
        q = np.ascontiguousarray(q, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        check("cam_box_set_state", lib().cam_box_set_state(self._h, _p(q), _p(p)))

    def get_state(self):
        q = np.empty((self.nbeads, self.natoms, 3))
        p = np.empty_like(q)
        check("cam_box_get_state", lib().cam_box_get_state(self._h, _p(q), _p(p)))
        return q, p

    def prepare(self, n):
        check("cam_box_prepare", lib().cam_box_prepare(self._h, n))

    def set_transrot(self, on=True):
        """transrot after every step, as dynamic.x does (src/verlet.f90:1299-1305)."""
        check("cam_box_set_transrot", lib().cam_box_set_transrot(self._h, int(on)))

    def set_nhc(self, nose_q):
        """Nose-Hoover chain thermostat with damping time nose_q (atomic time
        units; 0 turns it off).  Needs andersen_step 0."""
        check("cam_box_set_nhc", lib().cam_box_set_nhc(self._h, float(nose_q)))

    def run(self, n, sync=True):
        check("cam_box_run", lib().cam_box_run(self._h, n))
        if sync:
            return self.sync()

    def sync(self):
        ms = ctypes.c_double()
        check("cam_box_sync", lib().cam_box_sync(self._h, ctypes.byref(ms)))
        return ms.value

    def energies(self):
        ep = np.empty(self.nbeads)
        ek, es = ctypes.c_double(), ctypes.c_double()
        check("cam_box_energies", lib().cam_box_energies(self._h, _p(ep), ctypes.byref(ek), ctypes.byref(es)))
        return ep, ek.value, es.value
