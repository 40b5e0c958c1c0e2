"""The C-ABI library loads on a machine without a GPU and exports every
entry point declared in include/nbl_api.h (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "nabla-ai_amd" / "lib" / "libnabla_hip.so"


def declared():
    text = (ROOT / "include" / "nbl_api.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "nbl_run_step" in names and "nbl_poisson_solve" in names
    assert "nbl_octree_face_table" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_match_reference_flowconfig():
    if not LIB.exists():
        pytest.skip("HIP library not built")
    import nbl
    lib = nbl.load()
    c = nbl.FlowCfg()
    lib.nbl_flow_cfg_default(ctypes.byref(c))
    # FlowConfig{} defaults, core/include/nabla/flow/ns_solver.hpp:14-32
    assert (c.reynolds, c.ref_length, c.velocity_scale, c.density, c.cfl) == (100.0, 1.0, 1.0, 1.0, 0.8)
    assert (c.poisson_tol, c.poisson_max_iters, c.tol_continuity, c.tol_mass) == (1e-9, 5000, 1e-7, 1e-8)
    assert c.convection == nbl.CONV_CENTRAL and c.ref_area == 0.0
    o = nbl.SolverOpts()
    lib.nbl_solver_opts_default(ctypes.byref(o))
    assert (o.precision, o.pre_smooth, o.post_smooth) == (nbl.PREC_FP64, 2, 2)


def test_errors_are_reported_not_thrown():
    if not LIB.exists():
        pytest.skip("HIP library not built")
    import nbl
    lib = nbl.load()
    rc = lib.nbl_run_step(None, 0.0, None)
    assert rc == -1
    assert b"NULL" in lib.nbl_last_error()
