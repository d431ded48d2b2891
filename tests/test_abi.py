"""C ABI: the library loads and exports every entry point include/komd.h declares
(no compute calls, so this runs without a GPU)."""
import ctypes
import os

import komd_amd as K


def test_library_exports_every_declared_symbol():
    lib = K.load_library()
    names = K.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_wrapper_signatures_cover_the_header():
    names = set(K.declared_symbols())
    assert names == set(K._SIGNATURES), names ^ set(K._SIGNATURES)


def test_library_is_in_tree():
    assert os.path.dirname(K.LIB_PATH) == os.path.dirname(os.path.abspath(K.__file__))
    ctypes.CDLL(K.LIB_PATH)


def test_committed_surrogate_weights():
    """The shipped surrogate parameters have the documented size and are finite."""
    import numpy as np
    from komd_amd.train import DEFAULT_WEIGHTS, NPARAM
    p = np.load(DEFAULT_WEIGHTS)
    assert p.dtype == np.float32 and p.size == NPARAM and np.isfinite(p).all()
    assert NPARAM == 64 * 48 + 64 + 64 * 64 + 64 + 64 + 1 + 48 + 48


def test_max_neighbors_is_validated():
    """The neighbour-list capacity is fixed at 96 sites: komd_create accepts 0 or 96 and
    rejects any other value with KOMD_ERR_ARG (checked before any device call)."""
    import pytest
    from komd_amd import systems as S
    s = S.fcc(5, 0.8, 0.8, 1)
    for bad in (64, 128, -1):
        with pytest.raises(K.KomdError, match="max_neighbors") as ei:
            K.Context(s, max_neighbors=bad)
        assert ei.value.status == 1
