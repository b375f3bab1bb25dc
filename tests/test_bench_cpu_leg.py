"""CPU check of the measurement harness: bench.py's CPU-baseline leg (the
oracle's optimised build solving the C1 Wilson workload to convergence) runs
without a GPU and reports what the bench line documents: the build, the
host, the thread counts of its legs, and the same iteration count for every
thread count (the oracle's results do not depend on it)."""
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpu_baseline_leg_c1():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--config", "c1", "--cpu-only", "--cpu-full",
                        "--cpu-threads", "2,1"], capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    line = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    cb = line["cpu_baseline"]
    assert line["config"]["workload"] == "wilson_8^4_kappa0.12"
    assert cb["kind"] == "port" and cb["unit"] == "s/RHS"
    assert "-O3" in cb["build"] and "x86-64-v3" in cb["build"]
    assert cb["host"]["cpus_available"] >= 1 and cb["host"]["mem_gb"] > 0
    assert cb["cores"] == 2 and cb["value"] > 0
    legs = cb["other_legs"]
    assert len(legs) == 1 and legs[0]["cores"] == 1 and legs[0]["value"] > 0
    it2 = int(re.search(r"(\d+) iterations", cb["sample"]).group(1))
    it1 = int(re.search(r"(\d+) iterations", legs[0]["sample"]).group(1))
    assert it2 == it1
    res = float(re.search(r"true residual ([0-9.e+-]+)", cb["sample"]).group(1))
    assert res <= 1e-10
