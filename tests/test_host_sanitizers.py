"""The host parser, CPU codec and TIFF I/O under ThreadSanitizer and
AddressSanitizer + UndefinedBehaviorSanitizer: tests/host_sanitize.cpp built
with g++ against the product's host sources, run on archives large enough for
the threaded parse, and the TIFF export and import of the same pixels."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tilestream_amd", "csrc", "host")
SOURCES = [os.path.join(ROOT, "tests", "host_sanitize.cpp")] + [
    os.path.join(HOST, f) for f in ("parser.cpp", "cpu_codec.cpp", "sizing.cpp", "tiff.cpp")]


@pytest.mark.parametrize("sanitizer", ["thread", "address,undefined"])
def test_host_code_under_sanitizer(tmp_path, sanitizer):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_sanitize")
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread",
                            f"-fsanitize={sanitizer}", "-fno-sanitize-recover=all",
                            "-I" + os.path.join(ROOT, "include"), "-I" + HOST] + SOURCES + ["-o", exe, "-lz"],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-3000:])
    assert "Sanitizer" not in run.stderr, run.stderr[-3000:]
    assert run.stdout.strip().endswith("rc 0"), run.stdout
