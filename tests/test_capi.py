"""The C ABI library loads on the CPU and exports every entry point that
include/nekamd.h declares (no compute call needs a GPU here), and the
host-only entry points (basis, numbering) agree with the oracle."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nekamd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nk[hd]_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    from nek5000_amd import _lib

    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding covers exactly the declared surface
    assert set(_lib.SIGNATURES) == set(syms)


def test_error_path_without_init():
    from nek5000_amd import _lib

    lib = _lib.load()
    h = ctypes.c_int(-1)
    rc = lib.nkh_mesh_create(8, 4, ctypes.byref(h))
    assert rc == 3  # NKH_ERR_STATE: no nkh_init yet
    assert b"nkh_init" in lib.nkh_last_error()
    assert lib.nkh_zwgll(None, None, 8) == 1


def test_host_basis_matches_oracle(oracle):
    from nek5000_amd.mesh import gll

    for n in (2, 4, 8, 11, 16):
        z, w, D = gll(n)
        zo, wo = oracle.zwgll(n)
        assert np.array_equal(z, zo) and np.array_equal(w, wo)
        assert np.array_equal(D, oracle.dgll(zo))
