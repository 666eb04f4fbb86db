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
