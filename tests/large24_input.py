"""The 24^4, beta 6.0 input of the reference's 24^4 golden flow
(tests/golden/ref_large24.json), regenerated bit for bit by the CPU oracle's
seeded generator: cold start, 200 heatbath/overrelaxation iterations.

It takes minutes of CPU time, so the test session starts it in a background
process as soon as collection has selected a test that needs it (see
conftest.py); the test then waits for the file.  The process only runs the
oracle's C generator: it never touches the GPU.

    python -m tests.large24_input OUT    (writes OUT.part, renamed to OUT when done)
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

_proc = None
_path = None


def spec():
    with open(os.path.join(HERE, "golden", "ref_large24.json")) as f:
        return json.load(f)


def generate(out):
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle as O
    d = spec()
    g = d["generator"]
    shape = tuple(d["shape"])
    t0 = time.time()
    U = O.generate(shape, d["beta"], seed=g["seed"], start=g["start"], sweeps=0, or_per_hb=g["or_per_hb"])
    for k in range(0, g["therm_sweeps"], 20):
        O.generate(shape, d["beta"], seed=g["seed"], start="continue", sweeps=20, U=U, first_sweep=k,
                   or_per_hb=g["or_per_hb"])
        print("large24 input: %d of %d iterations, %.0f s" % (k + 20, g["therm_sweeps"], time.time() - t0), flush=True)
    U.tofile(out + ".part")
    os.replace(out + ".part", out)


def start():
    """Start the background generation (once per session)."""
    global _proc, _path
    if _proc is not None:
        return
    fd, _path = tempfile.mkstemp(prefix="lgf_large24_", suffix=".bin")
    os.close(fd)
    os.remove(_path)
    log = open(_path + ".log", "w")
    _proc = subprocess.Popen([sys.executable, "-m", "tests.large24_input", _path], cwd=ROOT, stdout=log,
                             stderr=subprocess.STDOUT)


def wait(timeout_s=900):
    """The generated links (file order, shape (V, 4, 18)) and their sha256."""
    import numpy as np
    if _proc is None:
        start()
    try:
        rc = _proc.wait(timeout=timeout_s)
    except subprocess.TimeoutExpired:
        _proc.kill()
        raise
    with open(_path + ".log") as f:
        log = f.read()
    if rc != 0:
        raise RuntimeError("generation of the 24^4 input failed:\n" + log)
    U = np.fromfile(_path, dtype=np.float64).reshape(-1, 4, 18)
    return U, hashlib.sha256(U.tobytes()).hexdigest(), log


def cleanup():
    if _proc is not None and _proc.poll() is None:
        _proc.kill()
        _proc.wait()
    for p in (_path, _path and _path + ".part", _path and _path + ".log"):
        if p and os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    generate(sys.argv[1])
