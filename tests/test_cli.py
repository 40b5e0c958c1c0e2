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


def test_bf16_precision_is_refused_exits_1(tmp_path):
    """"mixed-bf16" names a BF16 smoother-scratch form that is not built: the
    case is refused with an error instead of silently running another form."""
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"type": "lid-cavity", "resolution": 8, "precision": "mixed-bf16"}))
    r = run("run", str(p))
    assert r.returncode == 1 and "mixed-bf16" in r.stderr and "not implemented" in r.stderr


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


def read_appended_vtu(path):
    """Arrays of an appended raw VTU, in file order (UInt64 headers)."""
    b = Path(path).read_bytes()
    head, _, rest = b.partition(b'<AppendedData encoding="raw">\n   _')
    specs = []
    for part in head.split(b"<DataArray ")[1:]:
        attrs = dict(kv.split(b"=", 1) for kv in part.split(b"/>")[0].split())
        specs.append((attrs[b"type"].strip(b'"').decode(), int(attrs[b"offset"].strip(b'"'))))
    out = []
    for typ, off in specs:
        (n,) = struct.unpack_from("<Q", rest, off)
        dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[typ]
        out.append(np.frombuffer(rest, dt, n // np.dtype(dt).itemsize, off + 8).copy())
    assert rest.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_vtu_snapshots_match_the_reference_writer(tmp_path, orc, fmt):
    """Snapshots written on the run's writer thread: the ASCII files are byte
    for byte what the reference's writeGridVtu (case.cpp:175-241, restated in
    the oracle) prints for the same state (read back from the checkpoint of
    the same step); the binary files hold the same arrays at full precision."""
    case = {"type": "cube-channel", "resolution": 4, "reynolds": 500.0, "max_steps": 6, "snapshot_every": 3,
            "checkpoint_every": 3, "vtu_format": fmt, "run_dir": str(tmp_path / "r")}
    (tmp_path / "c.json").write_text(json.dumps(case))
    r = run("run", str(tmp_path / "c.json"))
    assert r.returncode == 0, r.stderr
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4, reynolds=500.0)
    for vtu, ckpt in (("snapshot_3.vtu", "checkpoint_3.ckpt"), ("snapshot_6.vtu", "checkpoint_6.ckpt"),
                      ("final.vtu", "final.ckpt")):
        hdr, (u, v, w, p) = read_ckpt(tmp_path / "r" / ckpt)
        o = orc.ns(g, cfg)
        o.set_state(u, v, w, p, hdr[3], hdr[4], hdr[5])
        if fmt == "ascii":
            o.write_vtu(tmp_path / "expected.vtu")
            assert (tmp_path / "r" / vtu).read_bytes() == (tmp_path / "expected.vtu").read_bytes(), vtu
        else:
            arrs = read_appended_vtu(tmp_path / "r" / vtu)
            n = g.cells
            cu, cv, cw = (o.cell_velocity(k) for k in range(3))
            i = np.arange(n) % g.nx
            j = (np.arange(n) // g.nx) % g.ny
            k = np.arange(n) // (g.nx * g.ny)
            x0, y0, z0 = g.origin[0] + i * g.dx, g.origin[1] + j * g.dy, g.origin[2] + k * g.dz
            x1, y1, z1 = x0 + g.dx, y0 + g.dy, z0 + g.dz
            pts = np.stack([x0, y0, z0, x1, y0, z0, x1, y1, z0, x0, y1, z0,
                            x0, y0, z1, x1, y0, z1, x1, y1, z1, x0, y1, z1], axis=1).reshape(-1)
            assert np.array_equal(arrs[0], pts)
            assert np.array_equal(arrs[1], np.arange(8 * n))
            assert np.array_equal(arrs[2], 8 * np.arange(1, n + 1))
            assert np.array_equal(arrs[3], np.full(n, 12, np.uint8))
            speed = np.sqrt(cu * cu + cv * cv + cw * cw)
            for got, want in zip(arrs[4:], (cu, cv, cw, p, speed, g.solid.astype(np.float64))):
                assert np.array_equal(got, want)
        o.close()
