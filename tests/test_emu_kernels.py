"""The library's HIP kernels run on the CPU kernel emulator (tests/emu), with
AddressSanitizer and exact-size device allocations, against the oracle.

Each case is a child process (the ASan runtime must be preloaded); an
out-of-bounds or use-after-free access in any kernel aborts it with ASan's
report.  These cases need no GPU: they check memory safety and results of
the same kernel sources that libtetra_hip.so is built from, at small sizes.
Case-name suffixes: _y preempts work-items at atomics, _t runs the
workgroups of each launch on four host threads at once, _r runs them in
reverse order.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"


@pytest.fixture(scope="module")
def emu_env():
    sys.path.insert(0, str(EMU))
    import build as emu_build

    emu_build.build()
    from oracle import binding

    binding.build()
    env = dict(os.environ)
    env["LD_PRELOAD"] = emu_build.asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    return env


CASES = [
    ("vv_cp", 5, 6, 1), ("vv_cp", 5, 6, 2), ("vv_cp", 5, 6, 3), ("vv_cp", 5, 6, 4),
    ("vv_cp", 7, 9, 2),
    # work-items preempted at atomics: compare-and-swap races in the bucket inserts
    ("vv_cp_y", 6, 7, 2), ("vv_cp_y", 6, 7, 4), ("vv_cp_y", 6, 7, 3),
    ("routes", 4, 1, 7), ("routes", 4, 2, 8),
    ("all_relations", 3, 4, 0), ("all_relations", 3, 4, 1),
    ("hubs", 11, 40), ("soup", 2, 60, 200),
    # cells drawn from windows of consecutive ids: many records per vertex,
    # short rows (window 16) or rows for the heavy kernel (window 40)
    ("window", 1, 100, 760, 16, 2), ("window", 3, 300, 2200, 40, 4),
    # TE, the vertex-lane hot path, then ET/TT/VF on the same mesh
    ("after_hot_path", 8, 5, 1), ("after_hot_path", 8, 5, 2),
    # workgroups of each launch run concurrently on host threads (their
    # global-memory accesses interleave), or in reverse order
    ("after_hot_path_t", 8, 5, 1), ("after_hot_path_r", 8, 5, 1),
    ("all_relations_t", 3, 4, 1), ("vv_cp_t", 5, 6, 2),
    # the no-allocation entry point: two calls, the second smaller
    ("noalloc", 6, 7, 4),
    ("split", 3, 4),
]


@pytest.fixture(scope="module")
def emu_results(emu_env):
    """All cases at once, a few child processes at a time."""
    from concurrent.futures import ThreadPoolExecutor

    def run(case):
        return subprocess.run([sys.executable, str(EMU / "run_case.py"), *map(str, case)],
                              env=emu_env, capture_output=True, text=True, timeout=600)

    with ThreadPoolExecutor(max(1, min(6, (os.cpu_count() or 2) - 1))) as ex:
        return dict(zip(CASES, ex.map(run, CASES)))


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_emulated(emu_results, case):
    r = emu_results[case]
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0 and "ok " in r.stdout, tail
