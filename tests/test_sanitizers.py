"""AddressSanitizer and UndefinedBehaviorSanitizer over the C oracle and the
product's host codec and tables: tests/sanitize/driver.cpp is built with
gcc / g++ -fsanitize=address,undefined (host code only) and run; any
sanitizer finding aborts it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CSRC = os.path.join(ROOT, "astc_implementation_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_oracle_and_host_codec_under_sanitizers(tmp_path):
    objs = []
    for src in ("or_tables.c", "or_color.c", "or_search.c", "or_pack.c"):
        out = str(tmp_path / (src + ".o"))
        subprocess.run(["gcc", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-I", ORACLE, *SAN, "-c",
                        os.path.join(ORACLE, src), "-o", out], check=True)
        objs.append(out)
    for src in ("astcg_tables.cpp", "astcg_codec.cpp"):
        out = str(tmp_path / (src + ".o"))
        subprocess.run(["g++", "-std=c++17", "-ffp-contract=off", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
                        *SAN, "-c", os.path.join(CSRC, src), "-o", out], check=True)
        objs.append(out)
    exe = str(tmp_path / "driver")
    subprocess.run(["g++", "-std=c++17", "-I", ORACLE, "-I", CSRC, "-I", os.path.join(ROOT, "include"), *SAN,
                    os.path.join(ROOT, "tests", "sanitize", "driver.cpp"), *objs, "-lm", "-lpthread", "-o", exe],
                   check=True)
    # verify_asan_link_order=0: the process may start with other preloaded
    # libraries ahead of the ASan runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "0 failed checks" in r.stdout
