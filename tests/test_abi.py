"""The C ABI: the library loads, exports every entry point include/caracal_amd.h
declares, and the ctypes structs match the C layout.  No compute calls."""
import ctypes
import os
import re
import subprocess

import pytest

import caracal_amd
from caracal_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "caracal_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cam_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds all of them
    assert set(names) == set(_abi.PROTOTYPES), set(names) ^ set(_abi.PROTOTYPES)


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "caracal_amd.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(cam_config), sizeof(cam_reaction),'
                   ' offsetof(cam_config, dt), offsetof(cam_config, seed), offsetof(cam_config, reaction));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split()
    c = _abi.Config
    assert [int(x) for x in out] == [ctypes.sizeof(c), ctypes.sizeof(_abi.Reaction), c.dt.offset, c.seed.offset,
                                     c.reaction.offset]


def test_version_and_error_text_need_no_gpu():
    assert caracal_amd.lib().cam_version().startswith(b"caracal_amd")
    assert isinstance(caracal_amd.lib().cam_last_error(), bytes)


def test_argument_errors_are_reported_not_aborted():
    # argument validation happens on the host before any HIP call
    lib = caracal_amd.lib()
    assert lib.cam_egrad_h3(None, 4, 1, None, None) == 1          # CAM_ERR_ARG
    assert b"natoms" in lib.cam_last_error()
    cfg = _abi.Config()
    cfg.surface = 7
    h = ctypes.c_void_p()
    assert lib.cam_engine_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert b"surface" in lib.cam_last_error()
    cfg.surface, cfg.natoms, cfg.nbeads, cfg.ntraj, cfg.dt = 1, 3, 10, 1, 1.0
    assert lib.cam_engine_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert b"nbeads" in lib.cam_last_error()


def test_fortran_module_binds_declared_names_only():
    """Every bind(C, name=...) of the Fortran host module is an entry point
    the header declares, and the module binds the engine and batch calls."""
    text = open(os.path.join(ROOT, "caracal_amd", "fortran", "caracal_amd_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(cam_\w+)"\)', text))
    assert bound <= set(declared_functions()), bound - set(declared_functions())
    for n in ("cam_engine_create", "cam_engine_run", "cam_engine_kappa", "cam_engine_autotune",
              "cam_batch_create", "cam_batch_run", "cam_batch_sync", "cam_water_create"):
        assert n in bound, n


def test_fortran_qmdff_desc_layout_matches_c(tmp_path):
    """The Fortran bind(C) cam_qmdff_desc has the C struct's size and field
    offsets (checked through the ctypes mirror of the header's struct)."""
    fc = "amdflang"
    mod_dir = os.path.join(ROOT, "caracal_amd", "fortran")
    if subprocess.run(["which", fc], capture_output=True).returncode != 0:
        pytest.skip("amdflang is not installed")
    if not os.path.exists(os.path.join(mod_dir, "caracal_amd_mod.mod")):
        pytest.fail("caracal_amd/fortran is not built")
    fields = [f[0] for f in _abi.QmdffDesc._fields_]
    lines = ["program layout", "use iso_c_binding", "use caracal_amd_mod",
             "type(cam_qmdff_desc), target :: d", "integer(c_intptr_t) :: b",
             "b = transfer(c_loc(d), b)", "print *, 'size', c_sizeof(d)"]
    for f in fields:
        lines.append(f"print *, '{f}', transfer(c_loc(d%{f}), b) - b")
    lines.append("end program layout")
    src = tmp_path / "layout.f90"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    # only the type is used: the module's .mod file, no object code
    subprocess.run([fc, "-I", mod_dir, str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = dict(zip(out[0::2], (int(v) for v in out[1::2])))
    assert got["size"] == ctypes.sizeof(_abi.QmdffDesc)
    for f in fields:
        assert got[f] == getattr(_abi.QmdffDesc, f).offset, f
