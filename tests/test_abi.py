"""libprh loads, exports every function include/prh.h declares, and its host
argument checks fail cleanly without a GPU.  CPU only (no kernel runs)."""
import ctypes as C
import os
import re

import pytest

from polyred_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "prh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert set(names) == set(abi.PROTOTYPES), names


def test_library_exports_every_symbol():
    lib = abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.prh_version().decode().startswith("libprh gfx950")


def test_library_is_gfx950_code_object():
    data = open(abi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_null_arguments_are_invalid():
    lib = abi.load()
    assert lib.prh_open(0, None) == abi.PRH_E_INVALID
    assert lib.prh_close(None) == abi.PRH_E_INVALID
    assert lib.prh_scene_upload(None, None, None) == abi.PRH_E_INVALID
    assert lib.prh_renderer_new(None, None, None, None, None) == abi.PRH_E_INVALID
    assert lib.prh_render(None) == abi.PRH_E_INVALID
    assert lib.prh_read_image(None, None, 0) == abi.PRH_E_INVALID
    assert lib.prh_read_buffer(None, 0, None, 0) == abi.PRH_E_INVALID
    assert lib.prh_last_stats(None, None) == abi.PRH_E_INVALID
    assert lib.prh_scene_free(None) == abi.PRH_E_INVALID
    assert lib.prh_renderer_free(None) == abi.PRH_E_INVALID
    assert b"null" in lib.prh_last_error()


def test_open_bad_ordinal():
    lib = abi.load()
    h = C.c_void_p()
    rc = lib.prh_open(9999, C.byref(h))
    assert rc in (abi.PRH_E_NODEVICE,)
    assert not h.value


def test_oracle_exports():
    import oracle
    L = oracle.lib()
    for n in ("oracle_render", "oracle_barycoord", "oracle_mat4_inv", "oracle_fragment_shader"):
        assert hasattr(L, n)
