"""Device job runner on the GPU: jobs running side by side produce the same
files as the nabla_solve subprocess the reference's LocalJobRunner launches
(checkpoints bit for bit), and a running job stops on cancel."""
import json
import subprocess
import sys
import threading
import time
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "nabla-ai_amd"))
import jobs  # noqa: E402

pytestmark = pytest.mark.gpu
EXE = ROOT / "nabla-ai_amd" / "bin" / "nabla_solve"

CASES = {
    "lid": {"type": "lid-cavity", "resolution": 24, "reynolds": 400.0, "max_steps": 12, "checkpoint_every": 6},
    "chan": {"type": "channel", "resolution": 12, "reynolds": 100.0, "max_steps": 10},
    "cube": {"type": "cube-channel", "resolution": 4, "reynolds": 500.0, "max_steps": 8, "snapshot_every": 4},
}


def test_side_by_side_jobs_match_the_cli(tmp_path):
    subs = []
    for name, c in CASES.items():
        for where in ("cli", "dev"):
            d = dict(c, run_dir=str(tmp_path / where / name))
            (tmp_path / f"{where}_{name}.json").write_text(json.dumps(d))
        r = subprocess.run([str(EXE), "run", str(tmp_path / f"cli_{name}.json")], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        subs.append((name, ["nabla_solve", "run", str(tmp_path / f"dev_{name}.json")], tmp_path / f"{name}.log"))
    done, ev = {}, threading.Event()

    def on_exit(rid, rc, cancelled):
        done[rid] = (rc, cancelled)
        if len(done) == len(subs):
            ev.set()

    runner = jobs.DeviceJobRunner(slots=2)
    try:
        for rid, argv, log in subs:
            runner.submit(rid, argv, on_exit, log)
        assert ev.wait(240)
    finally:
        runner.shutdown()
    assert done == {n: (0, False) for n in CASES}
    for name in CASES:
        a, b = tmp_path / "cli" / name, tmp_path / "dev" / name
        files = sorted(p.name for p in a.iterdir())
        assert files == sorted(p.name for p in b.iterdir())
        for f in files:
            if f.endswith(".ckpt") or f.endswith(".vtu"):
                assert (a / f).read_bytes() == (b / f).read_bytes(), (name, f)
        assert (a / "diagnostics.jsonl").read_bytes() == (b / "diagnostics.jsonl").read_bytes()
        assert "Nabla AI run" in (tmp_path / f"{name}.log").read_text()


def test_cancel_stops_a_running_job(tmp_path):
    d = {"type": "lid-cavity", "resolution": 32, "reynolds": 100.0, "max_steps": 100000,
         "run_dir": str(tmp_path / "long")}
    (tmp_path / "long.json").write_text(json.dumps(d))
    ev, out = threading.Event(), {}

    def on_exit(rid, rc, cancelled):
        out[rid] = (rc, cancelled)
        ev.set()

    runner = jobs.DeviceJobRunner(slots=1)
    try:
        runner.submit("long", ["nabla_solve", "run", str(tmp_path / "long.json")], on_exit, tmp_path / "long.log")
        t0 = time.time()
        while not (tmp_path / "long" / "diagnostics.jsonl").exists() and time.time() - t0 < 60:
            time.sleep(0.05)
        time.sleep(0.5)
        assert runner.cancel("long")
        assert ev.wait(60)
    finally:
        runner.shutdown()
    assert out["long"] == (-15, True)
    assert not (tmp_path / "long" / "final.ckpt").exists()
    assert "cancelled after step" in (tmp_path / "long.log").read_text()
