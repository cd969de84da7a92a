"""The C-ABI library loads and exports every function include/lantern_gpu.h
declares; host-only entry points and argument checks work without a GPU."""
import ctypes as C
import os
import re

import numpy as np

from lantern_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "lantern_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lt_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 30
    handle = C.CDLL(_capi.LIB_PATH)
    for n in names:
        assert hasattr(handle, n), n


def test_binding_covers_the_header():
    assert sorted(_capi.SIGNATURES) == _declared()


def test_generators_host_only():
    from lantern_amd import scenes
    p, n, i = scenes.geosphere(2.0, 3)
    assert p.shape == (1920, 3) and i.shape == (1280, 3)
    assert np.allclose(np.linalg.norm(p, axis=1), 2.0, atol=1e-5)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    gp, gn, guv, gi = scenes.grid(20.0, 10.0, 3, 5)
    assert gp.shape == (15, 3) and gi.shape == (16, 3)
    assert np.allclose(gn, [0, 1, 0])
    bp, bn, buv, bi = scenes.box(1.0, 2.0, 3.0)
    assert bp.shape == (24, 3) and bi.shape == (12, 3)


def test_argument_errors_are_reported():
    L = _capi.lib()
    out = C.c_void_p()
    assert L.lt_scene_create(None, C.byref(out)) == -1
    assert b"NULL" in L.lt_last_error()
    assert L.lt_film_download(None, None, None, None) == -1
    assert L.lt_render(None, None, None, None, None) == -1


def test_generators_match_oracle():
    import ctypes
    from lantern_amd import scenes
    from oracle import binding as ob
    for n in range(5):
        p, nr, idx = scenes.geosphere(1.0, n)
        nt = ctypes.c_uint32()
        nv = ob.lib().orc_geosphere(1.0, n, None, None, None, ctypes.byref(nt))
        P = np.zeros((nv, 3), np.float32)
        N = np.zeros((nv, 3), np.float32)
        I = np.zeros((nt.value, 3), np.uint32)
        ob.lib().orc_geosphere(1.0, n, ob.fp(P), ob.fp(N), ob.up(I), ctypes.byref(nt))
        assert np.array_equal(p, P) and np.array_equal(nr, N) and np.array_equal(idx, I)
    gp, gn, guv, gi = scenes.grid(7.0, 3.0, 4, 6)
    nt = ctypes.c_uint32()
    P = np.zeros((24, 3), np.float32); N = np.zeros((24, 3), np.float32)
    U = np.zeros((24, 2), np.float32); I = np.zeros((30, 3), np.uint32)
    ob.lib().orc_grid(7.0, 3.0, 4, 6, ob.fp(P), ob.fp(N), ob.fp(U), ob.up(I), ctypes.byref(nt))
    assert np.array_equal(gp, P) and np.array_equal(guv, U) and np.array_equal(gi, I)
    bp, bn, buv, bi = scenes.box(1.0, 2.0, 3.0)
    P = np.zeros((24, 3), np.float32); N = np.zeros((24, 3), np.float32)
    U = np.zeros((24, 2), np.float32); I = np.zeros((12, 3), np.uint32)
    ob.lib().orc_box(1.0, 2.0, 3.0, ob.fp(P), ob.fp(N), ob.fp(U), ob.up(I), ctypes.byref(nt))
    assert np.array_equal(bp, P) and np.array_equal(bn, N) and np.array_equal(buv, U) and np.array_equal(bi, I)
