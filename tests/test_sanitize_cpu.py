"""The CPU restatement and the host program under AddressSanitizer and
UndefinedBehaviorSanitizer (built by `make -C oracle sanitize` and
`make -C msc_project_amd lgf_asan`, host code only):

- oracle/_asan/orc_check calls every oracle entry point on small lattices
  (extents 2..6, so neighbour wraps and the 1x3 clover reach cross the
  lattice) and must finish without a sanitizer report;
- msc_project_amd/lgf_asan runs the host program's command-line parsing,
  path handling and error exits that need no GPU (bad arguments, a run
  without a usable GPU or input file) without a report."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
GOOD = ["4", "4", "6.0", "4", "4", "4", "4", "0", "0", "0", "0", "0", "0", "10", "0.01", "5"]
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:")


def clean(out):
    return not any(r in out.stdout + out.stderr for r in REPORTS)


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_asan", "orc_check")], capture_output=True, text=True,
                         timeout=600, env=ENV)
    assert out.returncode == 0 and "orc_check: ok" in out.stdout, out.stdout + out.stderr
    assert clean(out), out.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "msc_project_amd", "liblgf_flow.so")),
                    reason="liblgf_flow.so not built")
@pytest.mark.parametrize("args,status", [
    ([], {11}),
    (GOOD + ["{d}", "--bogus"], {11}),
    (GOOD + ["{d}", "--mode=nope"], {11}),
    (["4", "4", "6.0", "5", "4", "4", "4"] + GOOD[7:] + ["{d}"], {12}),
    (GOOD + ["{d}"], {13, 15}),  # no input file (13) or, on a CPU-only host, no GPU (15)
])
def test_host_program_under_sanitizers(tmp_path, args, status):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "msc_project_amd"), "lgf_asan"], check=True)
    d = str(tmp_path) + "/"
    out = subprocess.run([os.path.join(ROOT, "msc_project_amd", "lgf_asan"), *[a.format(d=d) for a in args]],
                         capture_output=True, text=True, timeout=120, env=ENV, stdin=subprocess.DEVNULL)
    assert out.returncode in status, (out.returncode, out.stderr)
    assert clean(out), out.stderr
