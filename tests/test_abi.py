"""The C ABI (include/ssn.h) against libssn.so and its ctypes binding, without a GPU.

Every function declared in the header is exported by the library and bound in
symplectic_stiefel_nearness_amd/_lib.py with the same number of parameters; the
host-only queries (version, padded order, status strings, workspace sizes) answer
correctly. No kernel is launched here.
"""
import ctypes
import re
from pathlib import Path

import pytest

from symplectic_stiefel_nearness_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "ssn.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(ssn_\w+)\s*\(([^)]*)\)\s*;", text):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libssn.so not built")
    return ctypes.CDLL(str(_lib.LIB_PATH))


def test_header_declares_the_bound_functions():
    decl = declared_functions()
    assert set(decl) == set(_lib.SIGNATURES), set(decl) ^ set(_lib.SIGNATURES)
    for name, nparams in decl.items():
        assert len(_lib.SIGNATURES[name][1]) == nparams, name


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_host_queries(lib):
    L = _lib.load()
    assert L.ssn_version().decode().startswith("ssn-gfx950")
    assert [L.ssn_padded_order(m) for m in (2, 16, 17, 96, 128)] == [16, 16, 32, 96, 128]
    assert L.ssn_status_string(_lib.ST_OK).decode() == "ok"
    for code in (_lib.ST_NO_DESCENT, _lib.ST_NEWTON_CAP, _lib.ST_NONFINITE, _lib.ST_EIGH):
        assert L.ssn_status_string(code).decode() not in ("", "ok")
    assert L.ssn_workspace_bytes(64, 4096, 96) > 64 * 8 * 96 * 96
    assert L.ssn_workspace_bytes(2, 100, 96) < L.ssn_workspace_bytes(4, 100, 96)
    assert L.ssn_workspace_bytes(1, 10, 3) == 0          # odd order: unsupported
    assert L.ssn_workspace_bytes(1, 10, 130) == 0        # above the largest order


def test_host_copy(lib):
    """ssn_host_copy is host-only: exact copies for odd sizes and thread counts, argument errors."""
    import numpy as np
    lib = _lib.load()                      # the bound entry points (size_t argument)
    rng = np.random.default_rng(0)
    for nbytes, threads in ((1, 1), (4097, 3), (1 << 20, 8), (12345, 64), (100, 200)):
        src = rng.integers(0, 256, nbytes, dtype=np.uint8)
        dst = np.zeros_like(src)
        assert lib.ssn_host_copy(dst.ctypes.data, src.ctypes.data, nbytes, threads) == 0
        assert np.array_equal(dst, src)
    assert lib.ssn_host_copy(None, None, 0, 4) == 0
    assert lib.ssn_host_copy(None, ctypes.c_void_p(1), 8, 4) == -1


def test_workspace_bytes_non_decreasing_in_batch(lib):
    """A workspace sized for a batch serves every smaller batch of the same shape: the Gram's chunk
    count falls as the batch grows, and the partial slabs are sized for the worst smaller batch.
    Also covers every chunk split of the CPU-in pipeline (sized by its biggest chunk)."""
    L = _lib.load()
    for n in (100, 2816, 4096, 8192):
        for m in (16, 96):
            sizes = [L.ssn_workspace_bytes(b, n, m) for b in range(1, 601)]
            assert all(a <= b for a, b in zip(sizes, sizes[1:])), (n, m)
            for batch in range(2, 401):
                nch = min(4, batch)
                bounds = [(batch * c // nch, batch * (c + 1) // nch) for c in range(nch)]
                biggest = max(b1 - b0 for b0, b1 in bounds)
                need = max(sizes[b1 - b0 - 1] for b0, b1 in bounds)
                assert sizes[biggest - 1] >= need, (n, m, batch)


def test_default_options_are_the_reference_constants(lib):
    L = _lib.load()
    o = _lib.SsnOptions()
    L.ssn_default_options(ctypes.byref(o))
    assert (o.max_newton, o.max_cg, o.max_backtrack, o.split, o.feas_tol, o.shared_device) == (100, 150, 40, 0, 1e-8, 0)
    from symplectic_stiefel_nearness_amd import dual
    d = dual._options(False)
    assert (d.max_newton, d.max_cg, d.max_backtrack, d.split, d.feas_tol, d.shared_device) == (100, 150, 40, 0, 1e-8, 0)
    assert dual._options(False, shared=True).shared_device == 1
    assert dual._options(True, dict(max_cg=0)).max_cg == 0 and dual._options(True).split == 1


def test_solve_opts_rejects_bad_options_before_any_launch(lib):
    """Argument errors come back before any HIP call (so this runs without a GPU)."""
    L = _lib.load()
    args = [None, 0, 0, 1, 16, 16, None, 0, 0, None, 0, 16, None, None, None, None, 0]
    assert L.ssn_solve_opts(None, None, *args) == _lib.SSN_ERR_ARGS
    for field, value in (("max_newton", -1), ("max_cg", -1), ("max_backtrack", 1101), ("split", 2),
                         ("feas_tol", float("nan")), ("feas_tol", -1.0), ("feas_tol", float("inf")),
                         ("shared_device", 2), ("shared_device", -1)):
        o = _lib.SsnOptions()
        L.ssn_default_options(ctypes.byref(o))
        setattr(o, field, value)
        assert L.ssn_solve_opts(None, ctypes.byref(o), *args) == _lib.SSN_ERR_ARGS, field
    o = _lib.SsnOptions()
    L.ssn_default_options(ctypes.byref(o))
    assert L.ssn_solve_opts(None, ctypes.byref(o), *args) == _lib.SSN_ERR_ARGS     # null T
