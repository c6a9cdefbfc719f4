"""include/vtx.h is a C header: a C99 program compiled with gcc against it and linked with
libvtx.so sees the struct layouts the ctypes binding declares and gets, through the host-only
mesh generator, the same constraints as the binding (no GPU needed)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from velvet_amd import _lib
from velvet_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def c_output(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = str(tmp_path_factory.mktemp("cabi") / "abi_check")
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_abi", "abi_check.c"), "-o", exe, "-L", libdir, "-lvtx",
           f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=60).stdout
    return out.splitlines()


def _field(lines, key):
    return [ln[len(key) + 1:] for ln in lines if ln.startswith(key + " ")]


def test_struct_layouts_match_the_binding(c_output):
    sizes = {k: int(v) for k, v in (ln.split() for ln in _field(c_output, "sizeof"))}
    assert sizes["vtx_params"] == C.sizeof(_lib.VtxParams)
    assert sizes["vtx_collider"] == C.sizeof(_lib.VtxCollider)
    assert sizes["vtx_tuning"] == C.sizeof(_lib.VtxTuning)
    assert sizes["vtx_server_config"] == C.sizeof(_lib.VtxServerConfig)
    assert sizes["vtx_server_stats"] == C.sizeof(_lib.VtxServerStats)
    assert _field(c_output, "defaults") == ["2 4 64 -9.800"]


def test_mesh_generator_through_c_matches_the_binding(c_output):
    assert _field(c_output, "rc") == ["0"]
    xyz = np.zeros((9, 3), np.float32)
    for a in range(3):
        for b in range(3):
            xyz[a * 3 + b] = (b, 0.5 if b == 2 else 0.0, a)
    tris = []
    for a in range(2):
        for b in range(2):
            i = a * 3 + b
            tris += [i, i + 1, i + 3, i + 3, i + 1, i + 4]
    st, sr, bd, br = S.mesh_constraints(xyz, np.array(tris, np.uint32))
    ns, nb = (int(v) for v in _field(c_output, "counts")[0].split())
    # 12 grid edges + 4 diagonals; interior edges (two triangles each): 16 - 8 boundary
    assert (ns, nb) == (16, 8) == (len(st), len(bd))
    c_st = [ln.split() for ln in _field(c_output, "stretch")]
    assert [(int(i), int(j)) for i, j, _ in c_st] == [tuple(e) for e in st.tolist()]
    assert np.array_equal(np.array([float(r) for _, _, r in c_st], np.float32), sr)
    c_bd = [ln.split() for ln in _field(c_output, "bend")]
    assert [tuple(int(v) for v in row[:4]) for row in c_bd] == [tuple(e) for e in bd.tolist()]
    assert np.array_equal(np.array([float(row[4]) for row in c_bd], np.float32), br)
    # the fold along the lifted column gives nonzero rest angles there, flat elsewhere
    assert np.any(br > 1e-3) and np.any(br == 0.0)
    assert _field(c_output, "small rc") == ["-5 error"]            # VTX_ECAPACITY
    assert "gfx950" in _field(c_output, "version")[0]
    assert _field(c_output, "abi") == [f"{_lib.ABI_VERSION} {_lib.ABI_VERSION}"]
