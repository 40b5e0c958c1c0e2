"""The in-process device job runner behind the backend's JobRunner interface
(backend/app/jobs.py:31-48): argument handling, error paths and cancellation
semantics that run before any device work (CPU); the runs themselves are in
test_gpu_jobs.py."""
import json
import sys
import threading
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "nabla-ai_amd"))
import jobs  # noqa: E402

HOST = ROOT / "nabla-ai_amd" / "lib" / "libnabla_host.so"


def collect(runner, subs):
    done, lock, ev = {}, threading.Lock(), threading.Event()

    def on_exit(rid, rc, cancelled):
        with lock:
            done[rid] = (rc, cancelled)
            if len(done) == len(subs):
                ev.set()

    for rid, argv, log in subs:
        runner.submit(rid, argv, on_exit, log)
    assert ev.wait(60)
    return done


def test_parse_argv_mirrors_the_cli():
    assert jobs.parse_argv(["run", "c.json"]) == ("c.json", None, None)
    assert jobs.parse_argv(["run", "c.json", "--restart", "k.ckpt"]) == ("c.json", "k.ckpt", None)
    assert jobs.parse_argv(["run"])[2].startswith("error: run requires")
    assert "unknown run option" in jobs.parse_argv(["run", "c.json", "--bogus"])[2]
    assert "missing value" in jobs.parse_argv(["run", "c.json", "--restart"])[2]
    assert jobs.parse_argv(["version"])[2].startswith("usage:")


def test_error_paths_and_interface(tmp_path):
    if not HOST.exists():
        pytest.skip("host library not built")
    (tmp_path / "bad.json").write_text(json.dumps({"type": "volcano"}))
    r = jobs.DeviceJobRunner(slots=2)
    try:
        done = collect(r, [("u", ["nabla_solve", "bogus"], tmp_path / "u.log"),
                           ("m", ["nabla_solve", "run", str(tmp_path / "nope.json")], tmp_path / "m.log"),
                           ("t", ["nabla_solve", "run", str(tmp_path / "bad.json")], tmp_path / "t.log")])
        assert done == {"u": (2, False), "m": (1, False), "t": (1, False)}
        assert (tmp_path / "u.log").read_text().startswith("usage:")
        assert (tmp_path / "m.log").read_text().startswith("error:")
        assert "unknown case type" in (tmp_path / "t.log").read_text()
        assert not r.is_active("m") and not r.cancel("m")
    finally:
        r.shutdown()


def test_duplicate_run_id_and_cancel_while_queued():
    """With the only slot held by a job whose on_exit blocks, a second job
    stays queued: a duplicate run id is refused, and cancelling it reports
    -15, cancelled, without running it."""
    r = jobs.DeviceJobRunner(slots=1)
    gate, seen = threading.Event(), {}
    try:
        r.submit("a", ["nabla_solve", "bogus"], lambda *a: gate.wait(10))
        r.submit("x", ["nabla_solve", "bogus"], lambda rid, rc, c: seen.update(x=(rc, c)))
        with pytest.raises(RuntimeError):
            r.submit("x", ["nabla_solve", "bogus"], lambda *a: None)
        assert r.is_active("x") and r.cancel("x")
    finally:
        gate.set()
        r.shutdown()
    assert seen["x"] == (-15, True)
