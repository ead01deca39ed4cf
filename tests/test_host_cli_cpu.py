"""CPU checks of the host program's command line (msc_project_amd/lgf): bad
arguments end with a message and exit status 10 + lgf_status before any GPU
work, where the reference's main would exit(EXIT_FAILURE) or wait on stdin
(LGF/LGF.cpp:20-37; the even-extent check of LGF/Lattice/Lattice.cpp:42-47)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LGF = os.path.join(ROOT, "msc_project_amd", "lgf")
GOOD = ["4", "4", "6.0", "4", "4", "4", "4", "0", "0", "0", "0", "0", "0", "10", "0.01", "5"]

pytestmark = pytest.mark.skipif(not os.path.exists(LGF), reason="host program not built")


def run(args, folder):
    return subprocess.run([LGF, *args], capture_output=True, text=True, timeout=60, stdin=subprocess.DEVNULL,
                          cwd=folder)


@pytest.mark.parametrize("args,status,message", [
    ([], 11, "missing arguments"),
    (GOOD, 11, "expected NrDims"),                       # no data folder
    (GOOD + ["{d}", "--bogus"], 11, "unknown option --bogus"),
    (GOOD + ["{d}", "--precision=fp16"], 11, "unknown option"),
    (GOOD + ["{d}", "--autotune=-1"], 11, "autotune step count"),
    (GOOD + ["{d}", "--checkpoint-every=-2"], 11, "checkpoint interval"),
    (GOOD + ["{d}", "--mode=nope"], 11, "unknown mode nope"),
    (["4", "4", "6.0", "5", "4", "4", "4"] + GOOD[7:] + ["{d}"], 12, "even number of lattice points"),
])
def test_bad_command_lines_fail_with_status_and_message(tmp_path, args, status, message):
    d = str(tmp_path) + "/"
    out = run([a.format(d=d) for a in args], str(tmp_path))
    assert out.returncode == status, (out.returncode, out.stderr)
    assert message in out.stderr
    assert os.listdir(tmp_path) == []  # nothing written
