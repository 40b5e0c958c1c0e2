"""nabla_solve CLI and the C++ host layer (case.json in; diagnostics.jsonl,
NBFCK01 checkpoints and VTU out), mirroring core/apps/nabla_solve/main.cpp and
core/src/flow/case.cpp."""
import json
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cases
import nbl

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "nabla-ai_amd" / "bin" / "nabla_solve"
KEYS = ["step", "t", "dt", "cfl", "momentum_residual", "continuity_residual", "divergence_max",
        "mass_error", "cd", "cl", "poisson_iters", "poisson_residual", "cells", "accepted"]


def run(*args, cwd=None):
    if not EXE.exists():
        pytest.skip("nabla_solve not built")
    return subprocess.run([str(EXE), *args], capture_output=True, text=True, cwd=cwd, timeout=300)


def test_usage_errors_exit_2():
    assert run().returncode == 2
    r = run("run")
    assert r.returncode == 2 and "requires a case JSON path" in r.stderr
    r = run("run", "x.json", "--bogus")
    assert r.returncode == 2 and "unknown run option" in r.stderr
    assert run("run", "x.json", "--restart").returncode == 2


def test_missing_case_file_exits_1(tmp_path):
    r = run("run", str(tmp_path / "nope.json"))
    assert r.returncode == 1 and r.stderr.startswith("error:")


def test_unknown_case_type_exits_1(tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"type": "volcano"}))
    r = run("run", str(p))
    assert r.returncode == 1 and "unknown case type" in r.stderr


def read_ckpt(path):
    b = Path(path).read_bytes()
    assert b[:8] == b"NBFCK01\0"
    nx, ny, nz, step = struct.unpack_from("<4i", b, 8)
    t, dt = struct.unpack_from("<2d", b, 24)
    off, arrs = 40, []
    for _ in range(4):
        (n,) = struct.unpack_from("<Q", b, off)
        off += 8
        arrs.append(np.frombuffer(b, np.float64, n, off).copy())
        off += 8 * n
    assert off == len(b)
    return (nx, ny, nz, step, t, dt), arrs


@pytest.mark.gpu
def test_lid_cavity_run_matches_oracle_and_restarts_bitwise(tmp_path, orc):
    case = {"type": "lid-cavity", "resolution": 16, "reynolds": 100.0, "max_steps": 10,
            "checkpoint_every": 5, "snapshot_every": 0, "run_dir": str(tmp_path / "a")}
    (tmp_path / "a.json").write_text(json.dumps(case))
    r = run("run", str(tmp_path / "a.json"))
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "a" / "diagnostics.jsonl").read_text().splitlines()
    meta = json.loads(lines[0])
    assert meta["event"] == "meta" and meta["solver"] == "nabla_solve"
    assert len(lines) == 11
    for q, ln in enumerate(lines[1:]):
        d = json.loads(ln)
        assert list(d.keys()) == KEYS and d["step"] == q + 1
    (hdr, (u, v, w, p)) = read_ckpt(tmp_path / "a" / "final.ckpt")
    assert hdr[:4] == (16, 16, 1, 10)
    assert (tmp_path / "a" / "final.vtu").read_text().startswith('<?xml version="1.0"?>')
    # the oracle (reference restatement) on the same case
    g, cfg = cases.lid_cavity(16, 100.0)
    o = orc.ns(g, cfg)
    o.initialize()
    for _ in range(10):
        o.step()
    ou, ov, ow, op, st, t, _dt = o.get_state()
    for a, b in zip((u, v, w, p), (ou, ov, ow, op)):
        assert np.linalg.norm(a - b) <= 1e-10 * max(np.linalg.norm(b), 1e-30)
    assert hdr[4] == pytest.approx(t, rel=1e-12)
    # restart from step 5 reproduces the straight run bit for bit
    case_b = dict(case, run_dir=str(tmp_path / "b"), restart_from=str(tmp_path / "a" / "checkpoint_5.ckpt"),
                  max_steps=5)
    (tmp_path / "b.json").write_text(json.dumps(case_b))
    r = run("run", str(tmp_path / "b.json"))
    assert r.returncode == 0, r.stderr
    (hdr_b, arrs_b) = read_ckpt(tmp_path / "b" / "final.ckpt")
    assert hdr_b == hdr
    for a, b in zip(arrs_b, (u, v, w, p)):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_bad_checkpoint_magic_exits_1(tmp_path):
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"NOTACKPT" + bytes(64))
    case = {"type": "lid-cavity", "resolution": 8, "max_steps": 1, "run_dir": str(tmp_path / "r")}
    (tmp_path / "c.json").write_text(json.dumps(case))
    r = run("run", str(tmp_path / "c.json"), "--restart", str(bad))
    assert r.returncode == 1 and "bad magic" in r.stderr
