"""The C ABI library: built, loadable, and exporting what include/nft.h declares."""
import ctypes
import os
import re

import pytest

from src_amd import nft

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "nft.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nft_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_wrapper_agree():
    assert set(declared()) == set(nft._SIGS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(nft.LIBPATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_library_is_gfx950():
    with open(nft.LIBPATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_init_rejects_bad_arguments_without_touching_a_gpu():
    L = nft.lib()
    h = ctypes.c_void_p()
    rc = L.nft_init(None, None, ctypes.byref(h))
    assert rc < 0 and b"null" in L.nft_last_error()


@pytest.mark.parametrize("field,value,words", [("precision", 3, b"precision"), ("precision", -1, b"precision"),
                                               ("itraj", 2, b"itraj"), ("precon", 4, b"preconditioner")])
def test_init_validates_parameters_without_touching_a_gpu(field, value, words):
    """nft_init checks its parameters before it looks for a device."""
    from tests.common import small_cfg
    L = nft.lib()
    p = nft.make_params(small_cfg(precon=1))
    setattr(p, field, value)
    mesh = nft.NftMesh()
    h = ctypes.c_void_p()
    rc = L.nft_init(ctypes.byref(mesh), ctypes.byref(p), ctypes.byref(h))
    assert rc < 0 and words in L.nft_last_error()
