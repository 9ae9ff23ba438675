"""The host C of the generator and the oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only): tests/sanitize_drive.c drives both over valid, corrupted, truncated and random input."""

import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.timeout(600)
def test_generator_and_oracle_are_clean_under_asan_and_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    exe = tmp_path / "drive"
    cmd = [gcc, "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           str(ROOT / "tests" / "sanitize_drive.c"), str(ROOT / "oracle" / "cpu_decode.c"),
           str(ROOT / "cubam_amd" / "synth" / "gen.c"), "-o", str(exe), "-lz", "-lpthread", "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=23", "UBSAN_OPTIONS": "print_stacktrace=1",
           "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=540)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "sanitized run ok" in r.stdout
