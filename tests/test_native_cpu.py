"""The C-ABI library on a machine without a GPU: it loads, exports every entry point the
public headers declare, and reports errors through its status codes."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from eden_amd import _native as N
from eden_amd import configs
from eden_amd.model import EdenHipError, Simulation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[\w\s\*]*?\b(edenhip_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_the_expected_entry_points():
    names = declared_functions()
    for n in ("edenhip_create", "edenhip_run", "edenhip_read_trace", "edenhip_read_events", "edenhip_reset",
              "edenhip_destroy", "edenhip_last_error", "edenhip_cable_setup"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = N.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in sorted(declared_functions()) if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)


def test_python_binding_covers_declared_symbols():
    assert declared_functions() <= set(N.EXPORTS)


def test_create_reports_errors_without_crashing():
    lib = N.load_library()
    h = C.c_void_p()
    # null descriptor: invalid, with a message
    rc = lib.edenhip_create(None, None, C.byref(h))
    assert rc == N.E_INVALID
    assert lib.edenhip_last_error(h)
    lib.edenhip_destroy(h)
    # a valid model on a machine without a GPU: device error, or (where a GPU is present) success
    m = configs.random_trees(2, sizes=[10, 20], t_final_ms=1.0)
    try:
        with Simulation(m) as s:
            assert s.total_rows == m.total_rows()
    except EdenHipError as e:
        assert "device" in str(e).lower() or "hip" in str(e).lower()


def test_cable_setup_rejects_malformed_trees():
    p = np.zeros((2, 4), np.float32)
    p[:, 3] = 1.0
    with pytest.raises(ValueError):
        N.cable_setup(np.array([-1, 1]), p, p, 1.0, 0.1)    # parent after child
    with pytest.raises(ValueError):
        N.cable_setup(np.array([0, 0]), p, p, 1.0, 0.1)     # no root


def test_model_row_count_follows_double_time_accumulation():
    m = configs.hh_population(n_cells=1, t_final_ms=100.0)
    assert m.total_rows() == 4001
    m = configs.random_trees(1, sizes=[5], t_final_ms=250.0)
    assert m.total_rows() == 10001
