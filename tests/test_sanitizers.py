"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code, on
the CPU.

- The oracle (oracle/Makefile `asan`, gcc) under its pins, step, adjoint,
  multigrid and output tests.
- libcfmi.so's host code (cajitafluids_amd/csrc/Makefile `asan`: the .cpp
  sources instrumented with -Xarch_host, linked with the device objects)
  under the host-only ABI, output and CLI tests.

Each run is a subprocess with the sanitizer runtime preloaded (libstdc++ too,
so that the runtime's C++ exception interceptor resolves); undefined
behaviour aborts the run (-fno-sanitize-recover).  Leak checking is off: the
Python interpreter itself does not free everything at exit.
"""
import glob
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_ASAN = os.path.join(REPO, "oracle", "_build", "libcf_oracle_asan.so")
CFMI_ASAN = os.path.join(REPO, "cajitafluids_amd", "csrc", "_build", "asan", "libcfmi_asan.so")


def _gcc_file(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True,
                          text=True, check=True).stdout.strip()


def _run(preload, extra_env, args):
    env = dict(os.environ)
    env.update(extra_env)
    env["LD_PRELOAD"] = " ".join(preload + [_gcc_file("libstdc++.so.6")])
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    r = subprocess.run([sys.executable] + args, cwd=REPO, env=env, capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    return r.stdout


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "asan"], check=True)
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(REPO, "cajitafluids_amd", "csrc"),
                    "asan"], check=True)
    return True


def test_oracle_under_asan_ubsan(built):
    out = _run([_gcc_file("libasan.so")], {"CF_ORACLE_LIBRARY": ORACLE_ASAN},
               ["-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                "tests/test_oracle_pins.py", "tests/test_oracle_step.py",
                "tests/test_oracle_adjoint.py", "tests/test_oracle_mg.py"])
    assert " passed" in out


def test_host_library_under_asan_ubsan(built):
    rt = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not rt:
        pytest.skip("clang ASan runtime not found")
    env = {"CFMI_LIBRARY": CFMI_ASAN}
    # the instrumented build is the one that loads
    probe = ("from cajitafluids_amd import _lib; _lib.load(); "
             "print(any('libcfmi_asan.so' in l for l in open('/proc/self/maps')))")
    assert _run(rt[:1], env, ["-c", probe]).strip() == "True"
    out = _run(rt[:1], env, ["-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                             "tests/test_cpu_abi.py", "tests/test_output.py",
                             "tests/test_cli.py"])
    assert " passed" in out
