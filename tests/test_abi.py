"""The C-ABI boundary: libcac_hip.so loads on a machine without a GPU,
exports every symbol include/cac_hip.h declares, and the product package
never reaches the oracle."""
import ctypes as C
import os
import re

import pytest

from cac_code_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = _lib.header_symbols()
    assert len(syms) >= 29
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_bindings_cover_header():
    # every declared function has a ctypes signature
    assert not [s for s in _lib.header_symbols() if s not in _lib.SIGNATURES]


def test_null_and_bad_arguments_rejected_without_gpu():
    L = _lib.lib()
    assert L.cac_ctx_destroy(None) == _lib.CAC_OK
    assert L.cac_set_box(None, None, None, None) == _lib.CAC_ERR_ARG
    assert L.cac_force(None, 0, 0, 0) == _lib.CAC_ERR_ARG
    assert L.cac_last_error(None) == b"null context"
    # bad table sizes are rejected before any device work
    nrho = C.c_int()
    assert L.cac_eam_tables_from_funcfl(2, 0.1, 2, 0.1, None, None, None, C.byref(nrho), None, None, None,
                                        None, None, None) == _lib.CAC_ERR_ARG


def test_context_creation_fails_loudly_without_device():
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    rc = _lib.lib().cac_ctx_create(0, C.byref(h))
    assert rc != _lib.CAC_OK and not h.value


def test_product_never_imports_oracle():
    pat = re.compile(r"^\s*(from|import)\s+oracle\b|liboracle|oracle/", re.M)
    for dirpath, _, files in os.walk(os.path.join(ROOT, "cac_code_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not pat.search(text), f


def test_surrogate_constants_match_header():
    import re
    from cac_code_amd import engine, surrogate
    text = open(_lib.HEADER).read()
    vals = {k: int(v) for k, v in re.findall(r"#define (CAC_SURR_\w+) (\d+)", text)}
    assert (engine.SURR_NFEAT, engine.SURR_NOUT, engine.SURR_NQUAD) == (
        vals["CAC_SURR_NFEAT"], vals["CAC_SURR_NOUT"], vals["CAC_SURR_NQUAD"])
    assert surrogate.NIN == engine.SURR_NFEAT and surrogate.NOUT == engine.SURR_NOUT
