"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

Builds a copy of libdgv_amd.so whose host code (grid geometry, the pair-block
layout, the A-array and solution files, every C-ABI argument check) is
instrumented -- device code unchanged -- and the oracle likewise, then runs the
host-side test files against them in a child process with the sanitizer
runtime preloaded.  Any invalid access, leak-free misuse or undefined
behaviour ends the child with a report; the test fails on it."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TESTS = ["test_aarray_io.py", "test_capi_host.py", "test_driver_io.py", "test_oracle_pins.py",
              "test_oracle_lists.py", "test_oracle_trunc.py", "test_oracle_mixed.py", "test_oracle_regime.py",
              "test_agen_host.py", "test_close_order.py", "test_golden_ref.py", "test_ref_io.py"]


def _asan_runtime():
    for p in sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")):
        return p
    return None


@pytest.mark.slow
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    rt = _asan_runtime()
    if rt is None or shutil.which("make") is None:
        pytest.skip("ROCm clang sanitizer runtime not found")
    sys.path.insert(0, ROOT)
    from dgv0d3v_direct_collision_amd import build
    out = str(tmp_path / "san")
    lib = build.build_sanitized(out)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"OUT={out}"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt + ((" " + env["LD_PRELOAD"]) if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["DGV_LIB_PATH"] = lib
    env["DGV_ORACLE_LIB"] = os.path.join(out, "libdgv_oracle.so")
    probe = ("import dgv0d3v_direct_collision_amd as d, oracle; d.load(); oracle.lib(); "
             "m = open('/proc/self/maps').read(); "
             f"assert {lib!r} in m and {env['DGV_ORACLE_LIB']!r} in m and 'libclang_rt.asan' in m, 'not loaded'")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    tests = [os.path.join(ROOT, "tests", t) for t in HOST_TESTS]
    r = subprocess.run([sys.executable, "-m", "pytest", *tests, "-q", "-x", "-m", "not gpu and not slow",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True)
    log = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, log
    assert "runtime error" not in r.stdout + r.stderr, log
    assert "AddressSanitizer" not in r.stdout + r.stderr, log
