"""ctypes mirror of include/caracal_amd.h.  The shared library is built
in-tree (caracal_amd/libcaracal_amd.so); loading fails loudly if it is not."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcaracal_amd.so")

CAM_OK = 0
CAM_SURFACE_H3, CAM_SURFACE_CH4H, CAM_SURFACE_OH3, CAM_SURFACE_BRH2, CAM_SURFACE_O3 = 1, 2, 3, 4, 5
NATOMS = {CAM_SURFACE_H3: 3, CAM_SURFACE_CH4H: 6, CAM_SURFACE_OH3: 4, CAM_SURFACE_BRH2: 3, CAM_SURFACE_O3: 3}
CAM_PREC_FP64, CAM_PREC_MIXED = 0, 1
CAM_NM_REFERENCE, CAM_NM_ORTHOGONAL = 0, 1
CAM_ENS_NVE, CAM_ENS_ANDERSEN, CAM_ENS_CHILD, CAM_ENS_PARENT, CAM_ENS_UMBRELLA, CAM_ENS_NHC = 0, 1, 2, 3, 4, 5
CAM_NMK_AUTO, CAM_NMK_VALU, CAM_NMK_MFMA, CAM_NMK_FOLDED = 0, 1, 2, 3
CAM_FF_PAIRS_LIST, CAM_FF_PAIRS_ALL = 0, 1
CAM_SCHED_AUTO, CAM_SCHED_FUSED, CAM_SCHED_SPLIT = 0, 1, 2
CAM_MAX_ATOMS, CAM_MAX_BONDS, CAM_MAX_REAC = 16, 4, 4


class Reaction(ctypes.Structure):
    _fields_ = [
        ("n_form", ctypes.c_int32), ("n_break", ctypes.c_int32),
        ("bond_form", (ctypes.c_int32 * 2) * CAM_MAX_BONDS),
        ("bond_break", (ctypes.c_int32 * 2) * CAM_MAX_BONDS),
        ("form_ref", ctypes.c_double * CAM_MAX_BONDS),
        ("break_ref", ctypes.c_double * CAM_MAX_BONDS),
        ("n_reac", ctypes.c_int32),
        ("reac_natoms", ctypes.c_int32 * CAM_MAX_REAC),
        ("reac_atoms", (ctypes.c_int32 * CAM_MAX_ATOMS) * CAM_MAX_REAC),
        ("r_inf", ctypes.c_double), ("xi_ideal", ctypes.c_double),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("surface", ctypes.c_int32), ("natoms", ctypes.c_int32), ("nbeads", ctypes.c_int32),
        ("ntraj", ctypes.c_int32), ("precision", ctypes.c_int32), ("nm_mode", ctypes.c_int32),
        ("ensemble", ctypes.c_int32), ("remove_com", ctypes.c_int32), ("andersen_step", ctypes.c_int32),
        ("steps_per_launch", ctypes.c_int32), ("nm_kernel", ctypes.c_int32), ("waves_per_eu", ctypes.c_int32),
        ("schedule", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("mass", ctypes.c_double * CAM_MAX_ATOMS),
        ("seed", ctypes.c_uint64),
        ("reaction", Reaction),
        ("nose_q", ctypes.c_double),
    ]


CAM_MAX_TRIALS = 32


class Tuning(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("chosen", ctypes.c_int32),
        ("nm_kernel", ctypes.c_int32 * CAM_MAX_TRIALS), ("steps_per_launch", ctypes.c_int32 * CAM_MAX_TRIALS),
        ("waves_per_eu", ctypes.c_int32 * CAM_MAX_TRIALS), ("split", ctypes.c_int32 * CAM_MAX_TRIALS),
        ("us_per_step", ctypes.c_double * CAM_MAX_TRIALS),
    ]


class QmdffDesc(ctypes.Structure):
    """cam_qmdff_desc: the reference's QMDFF module arrays."""
    _fields_ = [("natoms", ctypes.c_int32), ("nbond", ctypes.c_int32), ("nangl", ctypes.c_int32),
                ("ntors", ctypes.c_int32), ("nhb", ctypes.c_int32), ("nnci", ctypes.c_int32),
                ("at", ctypes.POINTER(ctypes.c_int32)), ("molnum", ctypes.POINTER(ctypes.c_double)),
                ("q", ctypes.POINTER(ctypes.c_double)), ("bond", ctypes.POINTER(ctypes.c_int32)),
                ("vbond", ctypes.POINTER(ctypes.c_double)), ("angl", ctypes.POINTER(ctypes.c_int32)),
                ("vangl", ctypes.POINTER(ctypes.c_double)), ("tors", ctypes.POINTER(ctypes.c_int32)),
                ("vtors", ctypes.POINTER(ctypes.c_double)), ("hb", ctypes.POINTER(ctypes.c_int32)),
                ("vhb", ctypes.POINTER(ctypes.c_double)), ("nci", ctypes.POINTER(ctypes.c_int32)),
                ("c6xy", ctypes.POINTER(ctypes.c_double)), ("r094", ctypes.POINTER(ctypes.c_double)),
                ("sr42", ctypes.POINTER(ctypes.c_double)), ("zab", ctypes.POINTER(ctypes.c_double)),
                ("r0ab", ctypes.POINTER(ctypes.c_double)), ("eps1", ctypes.c_double * 6),
                ("eps2", ctypes.c_double * 6), ("rad", ctypes.POINTER(ctypes.c_double)),
                ("scalehb", ctypes.POINTER(ctypes.c_double)), ("periodic", ctypes.c_int32),
                ("zahn", ctypes.c_int32), ("box", ctypes.c_double * 3), ("coul_cut", ctypes.c_double),
                ("vdw_cut", ctypes.c_double), ("cut_low", ctypes.c_double), ("zahn_a", ctypes.c_double),
                ("zahn_par", ctypes.c_double)]


class FFDesc(ctypes.Structure):
    _fields_ = [
        ("natoms", ctypes.c_int32), ("nbonds", ctypes.c_int32), ("nangles", ctypes.c_int32),
        ("type", ctypes.POINTER(ctypes.c_int32)), ("mol", ctypes.POINTER(ctypes.c_int32)),
        ("bonds", ctypes.POINTER(ctypes.c_int32)), ("angles", ctypes.POINTER(ctypes.c_int32)),
        ("box", ctypes.c_double), ("rcut", ctypes.c_double), ("zahn_a", ctypes.c_double),
        ("zahn_par", ctypes.c_double),
        ("sr42", (ctypes.c_double * 2) * 2), ("r0bj", (ctypes.c_double * 2) * 2),
        ("zrep", (ctypes.c_double * 2) * 2), ("arep", (ctypes.c_double * 2) * 2),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
PROTOTYPES = {
    "cam_last_error": (ctypes.c_char_p, []),
    "cam_version": (ctypes.c_char_p, []),
    "cam_egrad_h3": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "cam_egrad_ch4h": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "cam_egrad_oh3": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "cam_egrad_brh2": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "cam_egrad_o3": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "cam_egrad_h3_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "cam_egrad_ch4h_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "cam_egrad_oh3_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "cam_egrad_brh2_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "cam_egrad_o3_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "cam_egrad_noalloc": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "cam_engine_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "cam_engine_destroy": (ctypes.c_int, [_vp]),
    "cam_engine_set_state": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "cam_engine_get_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "cam_engine_get_forces": (ctypes.c_int, [_vp, _dp, _dp]),
    "cam_engine_draw_momenta": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int]),
    "cam_engine_constrain": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "cam_engine_begin_children": (ctypes.c_int, [_vp]),
    "cam_engine_begin_parent": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "cam_engine_set_windows": (ctypes.c_int, [_vp, _dp, _dp]),
    "cam_engine_begin_umbrella": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "cam_engine_umbrella_stats": (ctypes.c_int, [_vp, _dp, _dp]),
    "cam_engine_set_umbrella_histogram": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "cam_engine_umbrella_histogram": (ctypes.c_int, [_vp, _dp]),
    "cam_engine_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_engine_prepare": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_engine_sync": (ctypes.c_int, [_vp]),
    "cam_engine_kappa": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int]),
    "cam_engine_observables": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _ip]),
    "cam_engine_device_state": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                               ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "cam_engine_last_timing": (ctypes.c_int, [_vp, _dp, _ip, _ip]),
    "cam_engine_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(Tuning)]),
    "cam_water_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(_vp)]),
    "cam_batch_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int, ctypes.POINTER(_vp)]),
    "cam_batch_destroy": (ctypes.c_int, [_vp]),
    "cam_batch_engine": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "cam_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_batch_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_batch_sync": (ctypes.c_int, [_vp, _dp]),
    "cam_ff_create": (ctypes.c_int, [ctypes.POINTER(FFDesc), ctypes.POINTER(_vp)]),
    "cam_qmdff_create": (ctypes.c_int, [ctypes.POINTER(QmdffDesc), ctypes.POINTER(_vp)]),
    "cam_ff_destroy": (ctypes.c_int, [_vp]),
    "cam_ff_set_pair_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_ff_energy": (ctypes.c_int, [_vp, _dp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "cam_ff_forces": (ctypes.c_int, [_vp, _dp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "cam_box_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "cam_box_destroy": (ctypes.c_int, [_vp]),
    "cam_box_set_params": (ctypes.c_int, [_vp, _dp]),
    "cam_box_set_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "cam_box_get_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "cam_box_prepare": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_box_run": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_box_set_transrot": (ctypes.c_int, [_vp, ctypes.c_int]),
    "cam_box_set_nhc": (ctypes.c_int, [_vp, ctypes.c_double]),
    "cam_box_sync": (ctypes.c_int, [_vp, _dp]),
    "cam_box_energies": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "cam_box_device_state": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
}

_lib = None


def lib():
    """The loaded C ABI.  Raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class CamError(RuntimeError):
    def __init__(self, fn, code):
        msg = lib().cam_last_error().decode()
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def check(fn, code):
    if code != CAM_OK:
        raise CamError(fn, code)
