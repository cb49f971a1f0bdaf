"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(pnjl_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def _lib_path():
    from julia_relaxtime_amd import _lib, build
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    return _lib.LIB_PATH


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("pnjl_ctx_create", "pnjl_solve_fixedmu_batch", "pnjl_gap_residual_batch", "pnjl_autotune"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib_path())
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_host_only_entry_points():
    """Entry points that need no device: defaults and the host GL rule."""
    import numpy as np
    from julia_relaxtime_amd import _lib
    _lib_path()
    p = _lib.default_params()
    assert abs(p.Lambda - 602.3 / 197.327) < 1e-15
    assert abs(p.G - 1.835 / p.Lambda ** 2) < 1e-15
    o = _lib.default_opts()
    assert (o.xtol, o.ftol, o.residual_norm_max, o.max_iter) == (1e-9, 1e-9, 1e-6, 1000)
    x, w = _lib.gauleg(0.0, 10.0, 64)
    xr, wr = np.polynomial.legendre.leggauss(64)
    assert np.max(np.abs(x - (5 * xr + 5))) < 2e-14
    assert np.max(np.abs(w - 5 * wr)) < 1e-13
    with pytest.raises(TypeError):
        _lib.default_opts(not_an_option=1)
