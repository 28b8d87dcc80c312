"""C++ callers of include/fft_api.h (CPU).

tests/cxx/contract_driver.cpp is written the way the reference's own C++ callers are: the
default dimType of asdFftMakePlan1D, the seven- and eight-argument asdFftShift /
asdIfftShift of example/Ascend950/FFT/{fftshift,ifftshift}/*.cpp:117/:120, the five-argument
asdFftIstftMakePlan, out-parameters as references and AsdSip::ErrorType status names.  It
compiles against include/ unchanged, links libasdfft.so and runs the host-side contract
(no call reaches a device).  Its multi-threaded section creates, plans, queries and destroys
handles from 8 threads while others share one handle and one thread destroys a handle the
rest keep calling.  The same driver runs against the host library built with
AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer (the reference's
ASAN/UBSAN build switches: CMakeLists.txt:94, build.sh:1037-1038); a sanitizer report fails
the run (-fno-sanitize-recover, halt_on_error).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cxx", "contract_driver.cpp")
OUT = os.path.join(ROOT, "sip_amd", "_build", "tests")


def _run(cmd, env=None):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600, env=env)
    return r.returncode, r.stdout


def _quiet_env(**extra):
    env = dict(os.environ)
    env["ASDFFT_QUIET"] = "1"
    env.update(extra)
    return env


def test_reference_style_caller_against_shipped_library():
    from sip_amd import _native
    lib = _native.LIB_PATH
    if not os.path.exists(lib):
        pytest.skip("libasdfft.so is not built")
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "contract_driver")
    rc, out = _run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{ROOT}/include", DRIVER, "-o",
                    exe, f"-L{os.path.dirname(lib)}", "-lasdfft", f"-Wl,-rpath,{os.path.dirname(lib)}",
                    "-lpthread"])
    assert rc == 0, out
    rc, out = _run([exe, "8", "200"], env=_quiet_env())
    assert rc == 0, out
    assert "0 failed checks" in out


@pytest.mark.parametrize("sanitize,options", [
    ("address,undefined", {"ASAN_OPTIONS": "halt_on_error=1:detect_leaks=1",
                           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}),
    ("thread", {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"}),
])
def test_sanitized_host_library(sanitize, options):
    from sip_amd import build
    os.makedirs(OUT, exist_ok=True)
    exe = build.host_driver(DRIVER, os.path.join(OUT, "contract_driver_" + sanitize.replace(",", "_")), sanitize)
    rc, out = _run([exe, "8", "100"], env=_quiet_env(**options))
    assert rc == 0, out[-4000:]
    assert "0 failed checks" in out
    assert "Sanitizer" not in out, out[-4000:]
