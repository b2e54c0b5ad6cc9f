"""Host code under sanitizers, built with g++ against the product's host
sources.  The parser, CPU codec and TIFF I/O under ThreadSanitizer and
AddressSanitizer + UndefinedBehaviorSanitizer (tests/host_sanitize.cpp: archives
large enough for the threaded parse, TIFF export and import).  The planner, the
batched queue and the server under AddressSanitizer + UBSan on a CPU stand-in
for the HIP runtime whose kernel launchers check every table index a kernel
would read (tests/host_sanitize_plan.cpp, tests/host_hip_stub.cpp)."""
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
    # Each run writes its TIFF files under its own directory: parallel runs must not share them.
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               TMPDIR=str(tmp_path))
    run = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-3000:])
    assert "Sanitizer" not in run.stderr, run.stderr[-3000:]
    assert run.stdout.strip().endswith("rc 0"), run.stdout


def test_planner_under_address_sanitizer(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("HIP headers not available")
    exe = str(tmp_path / "host_sanitize_plan")
    sources = [os.path.join(ROOT, "tests", "host_sanitize_plan.cpp"), os.path.join(ROOT, "tests", "host_hip_stub.cpp")]
    sources += [os.path.join(HOST, f) for f in ("plan.cpp", "queue.cpp", "serve.cpp", "sizing.cpp", "parser.cpp",
                                                "cpu_codec.cpp")]
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread",
                            "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__",
                            "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + HOST] + sources +
                           ["-o", exe, "-lz"], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-3000:])
    assert "Sanitizer" not in run.stderr and "host stub" not in run.stderr, run.stderr[-3000:]
    assert run.stdout.strip().endswith("rc 0"), run.stdout


def test_parse_survives_thread_creation_failure(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_thread_failure")
    sources = [os.path.join(ROOT, "tests", "host_thread_failure.cpp")]
    sources += [os.path.join(HOST, f) for f in ("parser.cpp", "cpu_codec.cpp", "sizing.cpp")]
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I" + os.path.join(ROOT, "include"),
                            "-I" + HOST] + sources + ["-o", exe, "-ldl"], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-2000:])
    assert run.stdout.strip().endswith("rc 0"), run.stdout
