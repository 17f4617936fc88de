"""The C ABI libraries load without a GPU and export every function the headers in
include/ declare; argument errors come back as status codes."""
import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"pann.h": "libpann.so", "pann_tools.h": "libpann_tools.so"}


def declared_functions(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pann_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(header):
    lib = C.CDLL(os.path.join(REPO, "parlayann-gpu-busqueda_amd", HEADERS[header]))
    names = declared_functions(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{HEADERS[header]} lacks {missing}"


def test_python_binding_covers_the_header():
    from pann import _lib
    for header, sigs in (("pann.h", _lib.PANN_SIGNATURES), ("pann_tools.h", _lib.TOOLS_SIGNATURES)):
        assert set(declared_functions(header)) == set(sigs), header


def test_struct_layouts():
    from pann import _lib
    assert C.sizeof(_lib.QueryParams) == 32
    assert C.sizeof(_lib.SearchConfig) == 24
    assert C.sizeof(_lib.SearchOutputs) == 56


def test_argument_errors_are_status_codes():
    from pann import _lib
    L = _lib.lib()
    out = C.c_void_p()
    assert L.pann_vectors_create(None, 0, 0, C.byref(out)) == _lib.PANN_ERR_CAPACITY
    assert b"dim" in L.pann_last_error()
    assert L.pann_vectors_create(None, 1, 200, C.byref(out)) == _lib.PANN_ERR_CAPACITY
    assert L.pann_graph_create(0, 4, None, None, C.byref(out)) == _lib.PANN_ERR_ARG
    deg = (C.c_uint32 * 2)(3, 1)
    edges = (C.c_uint32 * 4)(1, 0, 1, 0)
    assert L.pann_graph_create(2, 2, deg, edges, C.byref(out)) == _lib.PANN_ERR_FORMAT
    assert L.pann_searcher_prepare(None, 1, None, C.byref(out)) == _lib.PANN_ERR_ARG
    assert L.pann_searcher_outputs(None, None) == _lib.PANN_ERR_ARG
    assert L.pann_copy_kernel(None, None, 3, None) == _lib.PANN_ERR_ARG
    # edges present by the degrees but a NULL edge pointer: refused before anything reads it
    deg = (C.c_uint32 * 3)(2, 1, 0)
    out = C.c_void_p()
    assert L.pann_graph_create(3, 2, deg, None, C.byref(out)) == _lib.PANN_ERR_ARG
    assert b"edges is NULL" in L.pann_last_error() and out.value is None


def test_device_count_without_device_is_an_error_code():
    from pann import _lib
    n = C.c_int(-1)
    code = _lib.lib().pann_device_count(C.byref(n))
    if code == 0:
        pytest.skip("a HIP device is visible here")
    assert code == _lib.PANN_ERR_HIP and n.value == 0
