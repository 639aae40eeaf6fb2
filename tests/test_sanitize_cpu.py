"""The host scene preparation (render_engine_amd/csrc/rc_scene.cpp) and the CPU oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (tests/native): the driver prepares the procedural scenes,
round-trips a scene file, runs the oracle's build, runner frames over an event, camera pass and adjoint,
checks the quantized wide nodes' boxes against the fp32 ones in exact arithmetic, then reads truncated and
corrupt scene files (CPU only; needs g++ with the sanitizer runtimes)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import render_engine_amd as rea

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def test_scene_preparation_and_oracle_under_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", HERE], check=True, timeout=600)
    good = str(tmp_path / "room.bin")
    rea.Scene.box_room(2, 400).save(good)
    b = np.fromfile(good, np.uint8)
    files = [good]
    for k, cut in enumerate((20, 90, len(b) // 3, len(b) - 9)):
        p = str(tmp_path / f"cut{k}.bin")
        b[:cut].tofile(p)
        files.append(p)
    m = b[60:76].view(np.int32)
    bn = 76 + 52 * int(m[0])   # after ST (36 B) and STN (16 B) per triangle
    nodes = b[bn:bn + 32 * int(m[1])].view(np.int32).reshape(-1, 8)
    inner = int(np.nonzero(nodes[:, 7] == 0)[0][0])
    cyc = b.copy()
    cyc[bn + 32 * inner + 12:bn + 32 * inner + 16] = np.array([0], np.int32).view(np.uint8)   # child = the root
    files.append(str(tmp_path / "cyclic.bin"))
    cyc.tofile(files[-1])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(HERE, "_build", "scene_check")] + files, cwd=HERE, env=env, capture_output=True, text=True,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert f"{good}: loaded" in out
    assert out.count(": loaded") == 1, out                  # every corrupt file refused
    assert "scene_check: 0 failed checks" in out
