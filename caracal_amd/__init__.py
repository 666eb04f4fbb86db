"""caracal_amd -- MI355X ring-polymer MD engine for Caracal's RPMD hot path.

Host-side mirror of the reference's plugin interface for this path, over the
C ABI of libcaracal_amd.so (HIP kernels for gfx950):

  egrad_h3(q) / egrad_ch4h(q)   the bead-batched analytic surfaces
                                (src/egrad_h3.f:29, src/egrad_ch4h.f:74)
  RPMDEngine                    the trajectory-batch replacement of verlet
                                (src/verlet.f90:65) and its force loop

Arrays follow the reference layout: a numpy array of shape (nbeads, natoms, 3)
(C order) is the Fortran array q(3,natoms,nbeads); a batch adds a leading
trajectory axis.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import (CAM_ENS_ANDERSEN, CAM_ENS_CHILD, CAM_ENS_NHC, CAM_ENS_NVE, CAM_ENS_PARENT,  # noqa: F401
                   CAM_ENS_UMBRELLA,
                   CAM_NM_ORTHOGONAL,
                   CAM_NM_REFERENCE, CAM_NMK_AUTO, CAM_NMK_FOLDED, CAM_NMK_MFMA, CAM_NMK_VALU, CAM_PREC_FP64,
                   CAM_SCHED_AUTO, CAM_SCHED_FUSED, CAM_SCHED_SPLIT, CAM_PREC_MIXED, CAM_SURFACE_CH4H, CAM_SURFACE_H3,
                   CAM_SURFACE_OH3, CAM_SURFACE_BRH2, CAM_SURFACE_O3,
                   CamError, check, lib)
from .units import (AMU_TO_ME, FS_TO_AU, KB_HARTREE, beta_of, mass_au)  # noqa: F401

_dp = ctypes.POINTER(ctypes.c_double)


def _as_c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _egrad(fn_name, natoms_req, q):
    q = _as_c(q)
    if q.ndim == 2:
        q = q[None]
    if q.ndim != 3 or q.shape[1] != natoms_req or q.shape[2] != 3:
        raise ValueError(f"q must have shape (nbeads, {natoms_req}, 3), got {q.shape}")
    nb = q.shape[0]
    v = np.empty(nb)
    g = np.empty_like(q)
    check(fn_name, getattr(lib(), fn_name)(_ptr(q), natoms_req, nb, _ptr(v), _ptr(g)))
    return v, g


def egrad_h3(q):
    """BKMP2 H3 surface for a stack of beads: q (nbeads, 3, 3) bohr ->
    (V (nbeads,) hartree, dV/dq (nbeads, 3, 3) hartree/bohr)."""
    return _egrad("cam_egrad_h3", 3, q)


def egrad_ch4h(q):
    """CH5 surface for a stack of beads: q (nbeads, 6, 3) bohr, atom order
    h4, C, h1, h2, h3, H_attacking -> (V, dV/dq)."""
    return _egrad("cam_egrad_ch4h", 6, q)


def egrad_oh3(q):
    """OH + H2 surface for a stack of beads: q (nbeads, 4, 3) bohr, atom order
    O, H1 (hydroxyl), H2, H3 -> (V, dV/dq)."""
    return _egrad("cam_egrad_oh3", 4, q)


def egrad_brh2(q):
    """BrH2 DIM-3C surface for a stack of beads: q (nbeads, 3, 3) bohr, atom
    order Ha, Br, Hc -> (V, dV/dq)."""
    return _egrad("cam_egrad_brh2", 3, q)


def egrad_o3(q):
    """O + O2 surface (Varga, Paukku, Truhlar) for a stack of beads: q
    (nbeads, 3, 3) bohr, three O atoms -> (V, dV/dq)."""
    return _egrad("cam_egrad_o3", 3, q)


def egrad_dev(surface, q_ptr, nbeads, v_ptr, g_ptr, precision=CAM_PREC_FP64, stream=None):
    """Device-pointer form (pointers are integers, e.g. torch tensor data_ptr())."""
    name = {CAM_SURFACE_H3: "cam_egrad_h3_dev", CAM_SURFACE_CH4H: "cam_egrad_ch4h_dev",
            CAM_SURFACE_OH3: "cam_egrad_oh3_dev", CAM_SURFACE_BRH2: "cam_egrad_brh2_dev",
            CAM_SURFACE_O3: "cam_egrad_o3_dev"}[surface]
    na = _abi.NATOMS[surface]
    check(name, getattr(lib(), name)(q_ptr, na, nbeads, v_ptr, g_ptr, precision, stream))


from .engine import RaggedBatch, RPMDEngine, Reaction, make_config  # noqa: E402,F401
