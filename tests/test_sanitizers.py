"""Host code under sanitizers (CPU only, no GPU).

Builds tests/sanitize/san_driver.cpp with the host-only C-ABI helpers
(h-blast_amd/csrc/hb_seq.cpp, hb_hdb.cpp, hb_balance.cpp, hb_seg.cpp, hb_cosched.cpp) and the
oracle sources twice -- with
-fsanitize=address,undefined and with -fsanitize=thread -- in a temporary
directory, runs them on random and malformed inputs, and fails on any report
or failed check.  Device code and the HIP runtime are not covered (GPU
sanitizers are not available for this target).
"""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "sanitize", "san_driver.cpp"),
       os.path.join(ROOT, "h-blast_amd", "csrc", "hb_seq.cpp"),
       os.path.join(ROOT, "h-blast_amd", "csrc", "hb_hdb.cpp"),
       os.path.join(ROOT, "h-blast_amd", "csrc", "hb_balance.cpp"),
       os.path.join(ROOT, "h-blast_amd", "csrc", "hb_seg.cpp"),
       os.path.join(ROOT, "h-blast_amd", "csrc", "hb_cosched.cpp")]
CSRC = [os.path.join(ROOT, "oracle", "hbo.c"), os.path.join(ROOT, "oracle", "hbo_gap.c"),
        os.path.join(ROOT, "oracle", "hbo_hdb.c"), os.path.join(ROOT, "oracle", "hbo_balance.c"),
        os.path.join(ROOT, "oracle", "hbo_seg.c")]
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")


def build_and_run(tmp, flags, args, env_extra):
    objs = []
    for c in CSRC:
        o = os.path.join(tmp, os.path.basename(c) + ".o")
        subprocess.run(["gcc", "-O1", "-g", "-std=c11", "-pthread", *flags, *INC, "-c", c,
                        "-o", o], check=True)
        objs.append(o)
    exe = os.path.join(tmp, "san")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", *flags, *INC, *SRC, *objs,
                    "-o", exe], check=True)
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=600)
    return r


def test_address_and_undefined_behaviour_sanitizers():
    with tempfile.TemporaryDirectory() as tmp:
        r = build_and_run(tmp, ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                                "-fno-omit-frame-pointer"], [tmp],
                          {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:"
                                           "verify_asan_link_order=0",
                           "UBSAN_OPTIONS": "print_stacktrace=1"})
        assert r.returncode == 0, r.stderr[-4000:]
        assert "ERROR" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]


def test_thread_sanitizer_on_threaded_oracle():
    with tempfile.TemporaryDirectory() as tmp:
        r = build_and_run(tmp, ["-fsanitize=thread"], ["threads"],
                          {"TSAN_OPTIONS": "halt_on_error=1"})
        assert r.returncode == 0, r.stderr[-4000:]
        assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
