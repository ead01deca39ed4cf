"""Writes the golden fixtures under tests/golden/ from the reference program.

Run in the build container only (it needs oracle/_ref/lgf_ref, which
`make -C oracle/ref` builds from the reference sources):

    python tests/golden/make_golden.py [small|config1|large24|resume|all]

Inputs are produced by the seeded generator of the CPU restatement
(oracle/lgf_oracle.c, orc_generate); their sha256 is recorded so that a test
can regenerate them bit for bit.  The reference program then flows them with
its own FlowSavedGaugeConfigurations / MeasureQdensityDistribution
(LGF/LGF.cpp:114-124,163-196) under `mpiexec -n 1`-equivalent singleton MPI,
printing E and Q with 17 significant digits.  Recorded per case:
  - t, E(t), Q(t) in the reference's print order (Q then E),
  - the reference's CSV files verbatim (6 significant digits, LGF default),
  - sha256 of the flowed configuration file and of the q(x) file, plus
    sum(q(x)) and a few q(x) samples,
  - the unimproved Q and improved E of the flowed configuration
    (MeasureFlowedGaugeConfigurations, LGF/LGF.cpp:228-255) and their CSVs.
The 4^4 case also stores its input configuration (147 KB) and the flowed
output configuration's first links so that format tests need no generator.
"""
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "lgf_ref")
IN_METHOD = "LHMC_3_of_0_330000"


def beta_str(beta):
    return ("%f" % beta).replace(".", "_")


def shape_str(shape):
    return "X".join(str(s) for s in shape)


def sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def run_ref(mode, shape, beta, c0, c1, steps, eps, every, folder, extra=()):
    args = [REF, mode, *extra, "4", "4", repr(beta), *[str(s) for s in shape], "0", "0", "0", "0",
            str(c0), str(c1), str(steps), repr(float(eps)), str(every), folder]
    t0 = time.time()
    out = subprocess.run(args, capture_output=True, text=True, check=True)
    wall = time.time() - t0
    Q = [float(x) for x in re.findall(r"Topological Charge = (\S+)", out.stdout)]
    E = [float(x) for x in re.findall(r"Energy Density\s*= (\S+)", out.stdout)]
    return Q, E, wall, out.stdout


def flow_case(name, shape, beta, configs, steps, eps, every, keep_input=False):
    """configs: list of float64 arrays (V,4,18) in file order."""
    folder = tempfile.mkdtemp(prefix="lgf_golden_") + "/"
    try:
        ens = os.path.join(folder, "ensembles", "su3", "beta" + beta_str(beta), shape_str(shape), IN_METHOD)
        os.makedirs(ens)
        inputs = []
        for i, U in enumerate(configs):
            p = os.path.join(ens, "torus_extdof4_%d.bin" % i)
            U.tofile(p)
            inputs.append(sha(p))
        Q, E, wall, _ = run_ref("flow", shape, beta, 0, len(configs) - 1, steps, eps, every, folder)
        nmeas = 1 + steps // every
        t = [0.0]
        tt = 0.0
        for s in range(1, steps + 1):
            tt += eps
            if s % every == 0:
                t.append(tt)
        res = {"name": name, "shape": list(shape), "beta": beta, "eps": eps, "steps": steps, "every": every,
               "input_sha256": inputs, "ref_wall_s_1rank": wall, "configs": []}
        obs = os.path.join(folder, "Observables")
        t_pick = eps * steps
        # q(x) on the flowed configurations (LGF/LGF.cpp:114-124)
        Qd, Ed, wall_q, _ = run_ref("qdensity", shape, beta, 0, len(configs) - 1, steps, eps, every, folder)
        # unimproved Q and improved E of the flowed configurations (LGF/LGF.cpp:228-255)
        Qm, Em, _, _ = run_ref("measured", shape, beta, 0, len(configs) - 1, steps, eps, every, folder)
        for i in range(len(configs)):
            c = {"t": t, "Q": Q[i * nmeas:(i + 1) * nmeas], "E": E[i * nmeas:(i + 1) * nmeas]}
            qcsv = os.path.join(obs, "Topological_Charge", "beta" + beta_str(beta), shape_str(shape), "GF",
                                "torus_extdof4_%d.csv" % i)
            ecsv = os.path.join(obs, "EnergyDensity", "beta" + beta_str(beta), shape_str(shape), "GF",
                                "torus_extdof4_%d.csv" % i)
            c["csv_Q"] = open(qcsv).read()
            c["csv_E"] = open(ecsv).read()
            flowed = os.path.join(folder, "ensembles", "su3", "beta" + beta_str(beta), shape_str(shape), "GF",
                                  "torus_extdof4_%d_Flowtime%f.bin" % (i, tt))
            c["flowed_name"] = os.path.basename(flowed)
            c["flowed_sha256"] = sha(flowed)
            qx = os.path.join(obs, "Qdensity", "beta" + beta_str(beta), shape_str(shape), "GF",
                              "torus_extdof1_%d_Flowtime%f.bin" % (i, t_pick))
            c["qx_name"] = os.path.basename(qx)
            c["qx_sha256"] = sha(qx)
            q = np.fromfile(qx, dtype=np.float64)
            c["qx_sum"] = float(q.sum())
            c["qx_abs_sum"] = float(np.abs(q).sum())
            idx = sorted(set([0, 1, len(q) // 3, len(q) // 2, len(q) - 1]))
            c["qx_samples"] = {str(k): float(q[k]) for k in idx}
            c["measured_Q_unimproved"] = Qm[i]
            c["measured_E_improved"] = Em[i]
            c["csv_Q_unimproved"] = open(os.path.join(obs, "UnImprovedTopological_Charge", "beta" + beta_str(beta),
                                                      shape_str(shape), "GF", "torus_extdof4_%d.csv" % i)).read()
            c["csv_E_improved"] = open(os.path.join(obs, "ImprovedEnergyDensity", "beta" + beta_str(beta),
                                                   shape_str(shape), "GF", "torus_extdof4_%d.csv" % i)).read()
            if keep_input:
                fl = np.fromfile(flowed, dtype=np.float64)
                c["flowed_first_link"] = fl[:18].tolist()
                c["flowed_last_link"] = fl[-18:].tolist()
            res["configs"].append(c)
        res["ref_wall_s_qdensity"] = wall_q
        return res
    finally:
        shutil.rmtree(folder, ignore_errors=True)


def case_small():
    shape, beta = (4, 4, 4, 4), 6.0
    U = O.generate(shape, beta, seed=4444, start="hot", sweeps=5)
    U.tofile(os.path.join(HERE, "cfg_4x4x4x4_hot5.bin"))
    res = flow_case("small_4x4x4x4", shape, beta, [U], steps=20, eps=0.02, every=5, keep_input=True)
    res["generator"] = {"seed": 4444, "start": "hot", "sweeps": 5, "or_per_hb": 4}
    res["input_file"] = "cfg_4x4x4x4_hot5.bin"
    return res


def config1_inputs():
    """4 configs, 8^4, beta=6.0: cold start, 200 sweeps (1 HB + 4 OR), save, +10 sweeps per save."""
    shape, beta, seed = (8, 8, 8, 8), 6.0, 20240601
    U = O.generate(shape, beta, seed=seed, start="cold", sweeps=200)
    out = [U.copy()]
    for k in range(3):
        O.generate(shape, beta, seed=seed, start="continue", sweeps=10, U=U, first_sweep=200 + 10 * k)
        out.append(U.copy())
    return shape, beta, seed, out


def case_config1():
    shape, beta, seed, cfgs = config1_inputs()
    res = flow_case("config1_8x8x8x8", shape, beta, cfgs, steps=100, eps=0.02, every=10)
    res["generator"] = {"seed": seed, "start": "cold", "therm_sweeps": 200, "sep_sweeps": 10, "or_per_hb": 4}
    res["plaquette"] = [O.avg_plaquette(shape, U) for U in cfgs]
    return res


def large24_input():
    shape, beta, seed = (24, 24, 24, 24), 6.0, 24242424
    U = O.generate(shape, beta, seed=seed, start="cold", sweeps=200)
    return shape, beta, seed, U


def case_large24(steps=50):
    shape, beta, seed, U = large24_input()
    res = flow_case("large_24x24x24x24", shape, beta, [U], steps=steps, eps=0.01, every=5)
    res["generator"] = {"seed": seed, "start": "cold", "therm_sweeps": 200, "or_per_hb": 4}
    res["plaquette"] = [O.avg_plaquette(shape, U)]
    return res


def case_resume():
    """ResumeFlowedConfiguration (LGF/LGF.cpp:197-226) on the 4^4 fixture: the
    reference flows it 10 RK3 steps (eps 0.02, every 5), then resumes the
    saved configuration at t = 0.2 for 10 more steps.  Recorded: the Q and E
    the resume prints (17 digits; it does not measure at the pickup time), the
    CSVs after both runs (the resume appends to them), and the configuration
    the resume saves."""
    shape, beta, eps, every = (4, 4, 4, 4), 6.0, 0.02, 5
    U = np.fromfile(os.path.join(HERE, "cfg_4x4x4x4_hot5.bin"), dtype=np.float64)
    folder = tempfile.mkdtemp(prefix="lgf_golden_") + "/"
    try:
        ens = os.path.join(folder, "ensembles", "su3", "beta" + beta_str(beta), shape_str(shape), IN_METHOD)
        os.makedirs(ens)
        U.tofile(os.path.join(ens, "torus_extdof4_0.bin"))
        run_ref("flow", shape, beta, 0, 0, 10, eps, every, folder)
        tt = 0.0
        for _ in range(10):
            tt += eps
        pickup = "%f" % tt  # std::to_string of the flow time, the saved file's name
        Q, E, wall, _ = run_ref("resume", shape, beta, 0, 0, 10, eps, every, folder, extra=(pickup,))
        t = []
        for s in range(1, 11):
            tt += eps
            if s % every == 0:
                t.append(tt)
        obs = os.path.join(folder, "Observables")
        gf = os.path.join(folder, "ensembles", "su3", "beta" + beta_str(beta), shape_str(shape), "GF")
        flowed = os.path.join(gf, "torus_extdof4_0_Flowtime%f.bin" % tt)
        fl = np.fromfile(flowed, dtype=np.float64)
        res = {"name": "resume_4x4x4x4", "shape": list(shape), "beta": beta, "eps": eps, "every": every,
               "first_steps": 10, "resume_steps": 10, "pickup": pickup, "input_file": "cfg_4x4x4x4_hot5.bin",
               "t": t, "Q": Q, "E": E,
               "csv_Q": open(os.path.join(obs, "Topological_Charge", "beta" + beta_str(beta), shape_str(shape),
                                          "GF", "torus_extdof4_0.csv")).read(),
               "csv_E": open(os.path.join(obs, "EnergyDensity", "beta" + beta_str(beta), shape_str(shape),
                                          "GF", "torus_extdof4_0.csv")).read(),
               "flowed_name": os.path.basename(flowed), "flowed_sha256": sha(flowed),
               "flowed_first_link": fl[:18].tolist(), "flowed_last_link": fl[-18:].tolist(),
               "ref_wall_s_1rank": wall}
        return res
    finally:
        shutil.rmtree(folder, ignore_errors=True)


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if not os.path.exists(REF):
        raise SystemExit("build the reference first: make -C oracle/ref")
    cases = {"small": case_small, "config1": case_config1, "large24": case_large24, "resume": case_resume}
    for k, fn in cases.items():
        if which in (k, "all"):
            t0 = time.time()
            res = fn()
            with open(os.path.join(HERE, "ref_%s.json" % k), "w") as f:
                json.dump(res, f, indent=1)
            print(k, "done in %.1f s" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
