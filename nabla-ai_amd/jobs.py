"""In-process job runner on the device (the backend's JobRunner seam,
backend/app/jobs.py:31-48).

The reference's LocalJobRunner (jobs.py:50-140) runs one `nabla_solve`
subprocess at a time from a FIFO queue.  DeviceJobRunner keeps the same
interface and the same file telemetry (each job still writes its run
directory: diagnostics.jsonl, snapshots, checkpoints, and its log) but runs
the jobs inside this process through `nbl_case_run` (include/nbl_host.h),
`slots` at a time: a queued job starts as soon as a slot frees up, and the
running jobs' device work overlaps on the GPU.  A job's results do not depend
on what runs beside it.  cancel() sets the job's flag, which the run reads
between steps (the subprocess runner's SIGTERM; on_exit then reports -15).
"""
from __future__ import annotations

import ctypes as C
import queue
import threading
from collections.abc import Callable
from dataclasses import dataclass, field
from pathlib import Path

import nbl

# (run_id, returncode, cancelled), invoked from a worker thread
OnExit = Callable[[str, int, bool], None]

_lib = None
_lib_lock = threading.Lock()


def _host_lib():
    global _lib
    with _lib_lock:
        if _lib is None:
            nbl.load()  # the device library first (the host library links it)
            lib = C.CDLL(str(Path(nbl.__file__).resolve().parent / "lib" / "libnabla_host.so"))
            lib.nbl_case_run.restype = C.c_int
            lib.nbl_case_run.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
            _lib = lib
        return _lib


@dataclass
class _Job:
    run_id: str
    argv: list[str]
    log_path: Path | None
    on_exit: OnExit
    flag: C.c_int = field(default_factory=lambda: C.c_int(0))


def parse_argv(argv):
    """`[exe] run <case.json> [--restart <ckpt>]` as nabla_solve parses it
    (main.cpp:131-175): (case, restart, None) or (None, None, usage error)."""
    args = list(argv)
    if "run" not in args:
        return None, None, "usage: nabla_solve run <case.json> [--restart <checkpoint>]"
    args = args[args.index("run") + 1:]
    case, restart, i = None, None, 0
    while i < len(args):
        a = args[i]
        if a == "--restart":
            if i + 1 >= len(args):
                return None, None, "error: missing value for --restart"
            restart = args[i + 1]
            i += 2
            continue
        if a.startswith("-"):
            return None, None, f"error: unknown run option '{a}'"
        case = a
        i += 1
    if case is None:
        return None, None, "error: run requires a case JSON path"
    return case, restart, None


class DeviceJobRunner:
    """JobRunner (submit / cancel / is_active / shutdown) running `nabla_solve
    run` jobs in this process on the GPU, `slots` concurrently."""

    def __init__(self, slots: int = 4) -> None:
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self._queue: queue.Queue[_Job | None] = queue.Queue()
        self._lock = threading.Lock()
        self._active: dict[str, _Job] = {}
        self._threads = [threading.Thread(target=self._worker, daemon=True, name=f"nabla-device-jobs-{k}")
                         for k in range(slots)]
        for t in self._threads:
            t.start()

    def submit(self, run_id: str, argv: list[str], on_exit: OnExit, log_path: Path | None = None) -> None:
        """Queue a job; on_exit fires exactly once per submit."""
        job = _Job(run_id, list(argv), Path(log_path) if log_path is not None else None, on_exit)
        with self._lock:
            if run_id in self._active:
                raise RuntimeError(f"run {run_id} already has an active job")
            self._active[run_id] = job
        self._queue.put(job)

    def cancel(self, run_id: str) -> bool:
        """Cancel a queued or running job; False if it is not active."""
        with self._lock:
            job = self._active.get(run_id)
            if job is None:
                return False
            job.flag.value = 1
        return True

    def is_active(self, run_id: str) -> bool:
        with self._lock:
            return run_id in self._active

    def shutdown(self) -> None:
        with self._lock:
            for job in self._active.values():
                job.flag.value = 1
        for _ in self._threads:
            self._queue.put(None)
        for t in self._threads:
            t.join(timeout=60.0)

    # -- worker threads ----------------------------------------------------
    def _worker(self) -> None:
        while True:
            job = self._queue.get()
            if job is None:
                return
            rc, cancelled = self._run_one(job)
            with self._lock:
                self._active.pop(job.run_id, None)
            job.on_exit(job.run_id, rc, cancelled)

    def _run_one(self, job: _Job) -> tuple[int, bool]:
        if job.flag.value:  # cancelled while queued
            return -15, True
        case, restart, usage = parse_argv(job.argv[1:] if job.argv else [])
        if usage is not None:
            self._log(job, usage + "\n")
            return 2, False
        buf = C.create_string_buffer(1 << 14)
        try:
            lib = _host_lib()
        except OSError as exc:
            self._log(job, f"failed to launch: {exc}\n")
            return 127, False
        rc = lib.nbl_case_run(case.encode(), restart.encode() if restart else None, C.byref(job.flag), buf,
                              C.sizeof(buf))
        self._log(job, buf.value.decode(errors="replace"))
        return rc, rc == -15

    @staticmethod
    def _log(job: _Job, text: str) -> None:
        if job.log_path is not None and text:
            with open(job.log_path, "ab") as f:
                f.write(text.encode())
