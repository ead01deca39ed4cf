"""Times the reference program (oracle/_ref/lgf_ref, built from the reference's
sources by oracle/ref/Makefile) in the build container, on a 24^4 lattice,
and writes the record profiles/cpu_reference_flow_24x24x24x24.json.

The reference never travels to the GPU host; this record is what the
benchmark reports as the reference's CPU throughput, beside the CPU oracle
timed live on the GPU host (bench.py's cpu_baseline) and in this container
(the bridge between the two hosts).

The harness's `timeflow` mode builds the objects FlowSavedGaugeConfigurations
(LGF/LGF.cpp:163-196) builds and times, on the slowest rank, n calls of
GradientFlow::flow() with no measurement due, and one measurement
(TopologicalCharge and EnergyDensity calculate(), what the flow runs every
measurement interval); loading and saving the configuration stay outside.
  flow only                  = 4 V n / flow time
  job, measuring every 5     = 4 V 5 / (5 flow time per step + measurement)
  (the 24^4 workload's interval)
at 1 MPI rank and at 8 (cuts 1,1,1,0: 2x2x2x1 blocks, the container's 8
cores), on an otherwise idle container.  The input is the CPU oracle's seeded
Haar-random 24^4 field (the cost does not depend on the values).

    python oracle/ref/time_reference.py [ranks ...]      (default: 1 8)
"""
import json
import os
import platform
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "lgf_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"
SHAPE = (24, 24, 24, 24)
BETA, EPS = 6.0, 0.01
OUT = os.path.join(ROOT, "profiles", "cpu_reference_flow_24x24x24x24.json")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run(folder, ranks, steps, cuts):
    args = [MPIEXEC, "-n", str(ranks), REF, "timeflow", "4", "4", repr(BETA), *map(str, SHAPE), *map(str, cuts),
            "0", "0", str(steps), repr(EPS), "1", folder]
    out = subprocess.run(args, capture_output=True, text=True, check=True)
    m = re.search(r"lgf_ref flow_s (\S+) steps (\d+) measure_s (\S+)", out.stderr)
    r = re.search(r"lgf_ref wall_s (\S+) ranks (\d+)", out.stderr)
    assert m and r and int(r.group(2)) == ranks and int(m.group(2)) == steps, out.stderr
    return float(m.group(1)), float(m.group(3)), float(r.group(1))


def main():
    ranks_list = [int(a) for a in sys.argv[1:]] or [1, 8]
    V = int(np.prod(SHAPE))
    folder = tempfile.mkdtemp(prefix="lgf_reftime_") + "/"
    rec = {"what": "reference program (LGF, built from its sources with g++ -O3 and MPICH) on a 24^4 lattice, "
                   "RK3 eps 0.01, timed in the build container: GradientFlow::flow() calls and one E + Q "
                   "measurement, slowest rank, file I/O excluded (oracle/ref/time_reference.py)",
           "host": {"cpu": cpu_model(), "logical_cpus": os.cpu_count(), "where": "build container"},
           "lattice": list(SHAPE), "runs": {}}
    try:
        d = os.path.join(folder, "ensembles", "su3", "beta6_000000", "24X24X24X24", "LHMC_3_of_0_330000")
        os.makedirs(d)
        O.generate(SHAPE, BETA, seed=1234, start="hot", sweeps=0).tofile(os.path.join(d, "torus_extdof4_0.bin"))
        for ranks in ranks_list:
            cuts = (0, 0, 0, 0) if ranks == 1 else (1, 1, 1, 0)
            steps = 4 if ranks == 1 else 12
            flow_s, meas_s, wall = run(folder, ranks, steps, cuts)
            flow = flow_s / steps
            rec["runs"][str(ranks)] = {
                "ranks": ranks, "cores": ranks, "cuts": list(cuts), "steps_timed": steps, "flow_s": flow_s,
                "flow_s_per_step": flow, "measurement_s": meas_s, "run_wall_s": wall,
                "flow_only_link_updates_per_s": 4 * V / flow,
                "job_link_updates_per_s_measure_every_5": 4 * V * 5 / (5 * flow + meas_s)}
            print(json.dumps(rec["runs"][str(ranks)]), flush=True)
    finally:
        shutil.rmtree(folder, ignore_errors=True)
    # the CPU oracle on the same host: the bridge to the GPU host, where
    # bench.py times it live (cpu_baseline)
    from bench import cpu_baseline
    rec["oracle_port"] = {}
    for threads in (1, os.cpu_count()):
        r = cpu_baseline(threads=threads)
        rec["oracle_port"]["threads_%d" % threads] = {"value": r["value"], "sample": r["sample"]}
        print("oracle port, %d threads: %.3g link-updates/s" % (threads, r["value"]), flush=True)
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
