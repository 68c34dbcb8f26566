"""The C-ABI library loads, exports every entry point include/ablate_fv.h declares, its struct layouts match the
Python binding, and host-side validation rejects bad input before any device work (no GPU needed)."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ablate_amd import _lib
from ablate_amd.mesh import BoxMesh

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "ablate_fv.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(abl_(?:fv|hbm)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), f"{n} declared in ablate_fv.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "binding signature table out of sync with the header"
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(abl_(?:fv|hbm)_\w+)$", out, flags=re.M))
    assert set(names) <= exported


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    fields = {
        "abl_fv_mesh_desc": [f[0] for f in _lib.MeshDesc._fields_],
        "abl_fv_physics_desc": [f[0] for f in _lib.PhysicsDesc._fields_],
        "abl_fv_bc_desc": [f[0] for f in _lib.BcDesc._fields_],
        "abl_fv_info": [f[0] for f in _lib.Info._fields_],
        "abl_fv_pgs_desc": [f[0] for f in _lib.PgsDesc._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for st, cls in (("abl_fv_mesh_desc", _lib.MeshDesc), ("abl_fv_physics_desc", _lib.PhysicsDesc),
                    ("abl_fv_bc_desc", _lib.BcDesc), ("abl_fv_info", _lib.Info), ("abl_fv_pgs_desc", _lib.PgsDesc)):
        assert int(got[st]) == ctypes.sizeof(cls), st
        for f in fields[st]:
            assert int(got[f"{st}.{f}"]) == getattr(cls, f).offset, f"{st}.{f}"


def _create(desc, phys):
    h = ctypes.c_void_p()
    rc = _lib.lib().abl_fv_create(ctypes.byref(desc), ctypes.byref(phys), _lib.FORM_FP64, ctypes.byref(h))
    return rc, _lib.lib().abl_fv_last_error(None).decode()


def _phys(**kw):
    base = dict(gamma=1.4, Rgas=287.0, n_species=0, flux_calculator=1, reconstruction=0, limiter=0,
                max_lim_grad=1e10, k=0.0, mu=0.0, diff=0.0, cfl=0.5)
    base.update(kw)
    return _lib.PhysicsDesc(**base)


def test_create_rejects_bad_physics():
    m = BoxMesh([4, 4], [0, 0], [1, 1])
    assert _create(m.desc, _phys(gamma=1.0))[0] != 0
    rc, msg = _create(m.desc, _phys(limiter=42))
    assert rc != 0 and "limiter" in msg
    rc, msg = _create(m.desc, _phys(flux_calculator=9))
    assert rc != 0 and "flux" in msg


def test_create_rejects_bad_mesh():
    m = BoxMesh([4, 4], [0, 0], [1, 1])
    fc = np.array(m.arrays.face_cells)
    fc[3, 1] = 10_000  # out of range
    d = _lib.MeshDesc.from_buffer_copy(m.desc)
    d.face_cells = fc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rc, msg = _create(d, _phys())
    assert rc != 0 and "out of range" in msg
    # a ghost cell range without boundary-condition slots
    mw = BoxMesh([4, 4], [0, 0], [1, 1], [False, True])
    d2 = _lib.MeshDesc.from_buffer_copy(mw.desc)
    d2.face_bc = None
    rc, msg = _create(d2, _phys())
    assert rc != 0 and "face_bc" in msg
    # negative volume
    vol = np.array(m.arrays.cell_volume)
    vol[5] = -1.0
    d3 = _lib.MeshDesc.from_buffer_copy(m.desc)
    d3.cell_volume = vol.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc, msg = _create(d3, _phys())
    assert rc != 0 and "volume" in msg


def test_null_arguments_return_status():
    L = _lib.lib()
    assert L.abl_fv_rhs(None, 0.0, None, None, 0, None) != 0
    assert b"null" in L.abl_fv_last_error(None)
    assert L.abl_fv_get_info(None, None) != 0
    assert L.abl_fv_mesh_box(3, None, None, None, None, 0, None) != 0


def test_hbm_probe_rejects_bad_arguments():
    """The bandwidth probe validates its arguments before touching the device."""
    lib = _lib.lib()
    g, ms = ctypes.c_double(), ctypes.c_double()
    assert lib.abl_hbm_probe(5, 1 << 30, 1, ctypes.byref(g), ctypes.byref(ms)) != 0
    assert "kind" in lib.abl_fv_last_error(None).decode()
    assert lib.abl_hbm_probe(0, 1024, 1, ctypes.byref(g), ctypes.byref(ms)) != 0  # below 1 MiB
    assert lib.abl_hbm_probe(1, 1 << 30, 0, ctypes.byref(g), ctypes.byref(ms)) != 0  # no repetitions
