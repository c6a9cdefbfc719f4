"""The C-ABI library loads and exports every symbol include/vtx.h declares (no GPU needed)."""
import ctypes as C

from velvet_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 30
    missing = [name for name in declared if not hasattr(L, name)]
    assert not missing, missing
    # the ctypes signature table covers the header exactly
    assert sorted(_lib._SIGS) == declared


def test_defaults_are_the_reference_defaults():
    p = _lib.VtxParams()
    _lib.lib().vtx_params_default(C.byref(p))
    assert (p.num_substeps, p.num_iterations, p.max_num_neighbors) == (2, 4, 64)
    assert abs(p.gravity[1] + 9.8) < 1e-6 and p.damping == 0.25 and p.relaxation_factor == 1.0
    assert abs(p.long_range_stretchiness - 1.2) < 1e-6 and abs(p.collision_margin - 0.06) < 1e-6
    assert abs(p.friction - 0.1) < 1e-6 and p.enable_self_collision == 1 and p.interleaved_hash == 3
    assert p.hash_cell_size_scalar == 1.5 and p.particle_diameter_scalar == 1.5
    assert b"gfx950" in _lib.lib().vtx_version()


def test_abi_version_matches_header():
    import re
    txt = open(_lib.HEADER).read()
    assert int(re.search(r"#define VTX_ABI_VERSION (\d+)", txt).group(1)) == _lib.ABI_VERSION
    assert _lib.lib().vtx_abi_version() == _lib.ABI_VERSION


def test_struct_sizes_match_header():
    assert C.sizeof(_lib.VtxParams) == 4 * 21
    assert C.sizeof(_lib.VtxCollider) == 4 * (1 + 3 + 3 + 16 + 16 + 1)


def test_hash_objects_rejects_bad_params_before_any_device_work():
    # validation happens on the host: these calls return errors without touching the (fake)
    # pointers, so they run without a GPU
    L = _lib.lib()
    fake = C.c_void_p(16)
    base = dict(num_objects=10, max_num_neighbors=64, cell_spacing=0.1, cell_spacing2=0.01, table_size=20,
                particle_diameter2=0.0025)
    for bad, code in ((dict(max_num_neighbors=0), -1), (dict(max_num_neighbors=65), -1), (dict(table_size=0), -1),
                      (dict(table_size=1 << 28), -1), (dict(num_objects=1 << 25), -5), (dict(cell_spacing=0.0), -1)):
        prm = _lib.VtxHashParams(**{**base, **bad})
        rc = L.vtx_hash_objects(fake, fake, fake, fake, fake, fake, fake, C.byref(prm), None)
        assert rc == code, (bad, rc, L.vtx_last_error())
    prm = _lib.VtxHashParams(**base)
    assert L.vtx_hash_objects(fake, fake, None, fake, fake, fake, fake, C.byref(prm), None) == -1
    assert L.vtx_hash_objects(fake, fake, fake, fake, fake, fake, fake, None, None) == -1


def test_hash_objects_noalloc_rejects_oversized_calls_before_any_device_work():
    # the fixed workspace holds 40.5 B per object plus 8 B per cell (64 MiB): larger calls are
    # refused on the host, so these run without a GPU
    L = _lib.lib()
    fake = C.c_void_p(16)
    base = dict(num_objects=1 << 21, max_num_neighbors=64, cell_spacing=0.1, cell_spacing2=0.01,
                table_size=1 << 22, particle_diameter2=0.0025)
    prm = _lib.VtxHashParams(**base)
    assert L.vtx_hash_objects_noalloc(fake, fake, fake, fake, fake, fake, fake, C.byref(prm), None) == -5
    assert b"scratch" in L.vtx_last_error()
    prm = _lib.VtxHashParams(**{**base, "num_objects": 1 << 20, "table_size": 1 << 24})
    assert L.vtx_hash_objects_noalloc(fake, fake, fake, fake, fake, fake, fake, C.byref(prm), None) == -5
    prm = _lib.VtxHashParams(**{**base, "max_num_neighbors": 0})
    assert L.vtx_hash_objects_noalloc(fake, fake, fake, fake, fake, fake, fake, C.byref(prm), None) == -1
