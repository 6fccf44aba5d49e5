"""The library's host C under AddressSanitizer + UndefinedBehaviorSanitizer
and ThreadSanitizer (no GPU): the threaded periodic catalogue reader
(capow_ascii.c), the threaded survey reader, n(z), FKP, sums and packing
(capow_survey_ingest.c) and the output math and writer (capow_host.c),
driven by tests/sanitize_driver.c on generated files with 1 and 8 threads
and through their failure paths; and the drop-in host program
(capow_amd/host/capow_main.c, built with ASan + UBSan against the normal
library) through the failure paths of tests/test_host.py, which exit before
any GPU work.  Any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "capow_amd", "csrc")
INC = os.path.join(ROOT, "include")
HOST_SRCS = [os.path.join(CSRC, f) for f in ("capow_ascii.c", "capow_survey_ingest.c",
                                              "capow_host.c")]
DRIVER = os.path.join(ROOT, "tests", "sanitize_driver.c")
SAN_ENV = dict(ASAN_OPTIONS="halt_on_error=1:abort_on_error=0:exitcode=23:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=24",
               TSAN_OPTIONS="halt_on_error=1:exitcode=25")


def _gcc():
    return shutil.which("gcc")


def _build(tmp_path, flags, name):
    exe = str(tmp_path / name)
    cmd = [_gcc(), "-O1", "-g", "-std=gnu11", "-pthread", *flags, "-I", INC, DRIVER, *HOST_SRCS,
           "-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer runtime not available: %s" % r.stderr.strip()[-300:])
    return exe


def _run(exe, tmp_path):
    env = dict(os.environ)
    env.update(SAN_ENV)
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]


@pytest.mark.skipif(_gcc() is None, reason="gcc not found")
def test_host_c_under_asan_and_ubsan(tmp_path):
    exe = _build(tmp_path, ["-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                            "-fno-sanitize-recover=undefined"], "drv_asan")
    _run(exe, tmp_path)


@pytest.mark.skipif(_gcc() is None, reason="gcc not found")
def test_host_c_under_tsan(tmp_path):
    exe = _build(tmp_path, ["-fsanitize=thread"], "drv_tsan")
    _run(exe, tmp_path)


@pytest.mark.skipif(_gcc() is None, reason="gcc not found")
def test_host_program_failure_paths_under_asan(tmp_path):
    """The drop-in program built with ASan + UBSan (host code; the library is
    the normal build) on the failure paths that exit before the GPU."""
    lib = os.path.join(ROOT, "capow_amd", "libcapow.so")
    if not os.path.exists(lib):
        pytest.skip("libcapow.so not built")
    exe = str(tmp_path / "capow_asan")
    host = [os.path.join(ROOT, "capow_amd", "host", f)
            for f in sorted(os.listdir(os.path.join(ROOT, "capow_amd", "host"))) if f.endswith(".c")]
    cmd = [_gcc(), "-O1", "-g", "-std=gnu11", "-fsanitize=address,undefined",
           "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-I", INC, *host, "-o", exe,
           "-L", os.path.dirname(lib), "-lcapow", "-Wl,-rpath," + os.path.dirname(lib), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer runtime not available: %s" % r.stderr.strip()[-300:])
    from tests.hostfiles import write_catalogue, write_params
    env = dict(os.environ)
    env.update(SAN_ENV)
    # the HIP runtime the library links allocates at load time and is not
    # instrumented: leaks are not checked for this program
    env["ASAN_OPTIONS"] = env["ASAN_OPTIONS"].replace("detect_leaks=1", "detect_leaks=0")
    env.pop("LD_PRELOAD", None)

    def run(*args):
        r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=120)
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
        return r.returncode, r.stdout

    rc, out = run("a", "b")
    assert rc == 1 and "Input parameters not found" in out
    rc, out = run(str(tmp_path / "nope.dat"), "d", "r", "o")
    assert rc == 1 and "not found" in out
    pf = write_params(str(tmp_path / "p1.dat"), drop=("NK",))
    rc, out = run(pf, "d", "r", "o")
    assert rc == 1 and "I miss a value for tag 'NK'" in out
    pf = write_params(str(tmp_path / "p2.dat"), extra=(("NX", 64), ("Bogus", 3)))
    rc, out = run(pf, "d", "r", "o")
    assert rc == 1 and "Tag 'Bogus' not allowed" in out
    pf = write_params(str(tmp_path / "p3.dat"))
    cat = str(tmp_path / "bad.txt")
    with open(cat, "w") as f:
        f.write("1 2 3 4 5 6\n1 2 oops\n")
    rc, out = run(pf, cat, "r", str(tmp_path / "o.txt"))
    assert rc == 1 and "Fatal Error" in out
    del write_catalogue
