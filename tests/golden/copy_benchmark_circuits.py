"""Copies the shared random circuits of the reference's sampling benchmark
(SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/circuits/circuit_NXX.txt,
N = 2, 4, ..., 20; 1000 gates each, written by generate_circuits.py with
seed 42) into tests/golden/circuits/ unchanged, and the reference's own
published CPU timings for the three benchmarks this package is compared on
into tests/golden/reference_timings.json:

* random-circuit sampling, ComplexF64 with gate fusion
  (benchmark_random_circuit_sampling/timings_smoq_Float64.txt);
* the 4-layer circuit + 3N local + 3(N-1) neighbour observables
  (benchmark_expectation_values/timing_smoq.txt);
* MCWF with first-order Trotter and σ_z dephasing, 1000 trajectories, and
  the same model as a density matrix with Trotter gates and with the exact
  propagator (scripts_quantum_simulations/demo_lidbladian_dm_vs_mcwf_exact_vs_trotter/
  output.log, TIMING SUMMARY "MCWF+Trotter", "DM+Trotter", "DM+Exact" rows).

The files are data: gate lists and timings.  Run once:
python tests/golden/copy_benchmark_circuits.py
"""
from __future__ import annotations

import json
import os
import re
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
BENCH = os.path.join(REF, "SmoQ.jl_benchmarks")


def read_two_columns(path: str) -> dict:
    out = {}
    for line in open(path, encoding="utf-8"):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        n, t = line.split()[:2]
        out[n] = float(t)
    return out


def demo_timings(path: str, method: str) -> dict:
    """The TIMING SUMMARY rows of one method, per N (skipped sizes omitted)."""
    text = open(path, encoding="utf-8").read()
    out = {}
    for block in re.split(r"\n={10,}\n\s+N = ", text)[1:]:
        n = block.split()[0]
        m = re.search(re.escape(method) + r"\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)", block)
        if m:
            out[n] = float(m.group(1))
    return out


def main() -> None:
    src = os.path.join(BENCH, "benchmark_random_circuit_sampling", "circuits")
    dst = os.path.join(HERE, "circuits")
    os.makedirs(dst, exist_ok=True)
    for n in range(2, 22, 2):
        name = f"circuit_N{n:02d}.txt"
        shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
    timings = {
        "random_circuit_sampling": {
            "source": "SmoQ.jl_benchmarks/benchmark_random_circuit_sampling/"
                      "timings_smoq_Float64.txt",
            "workload": "load + fuse 1000 gates, simulate once from |0...0>, cumsum(abs2), "
                        "100 searchsortedfirst samples; ComplexF64",
            "seconds": read_two_columns(os.path.join(
                BENCH, "benchmark_random_circuit_sampling", "timings_smoq_Float64.txt")),
        },
        "expectation_values": {
            "source": "SmoQ.jl_benchmarks/benchmark_expectation_values/timing_smoq.txt",
            "workload": "4 layers [H even, Ry all, CNOT odd, S odd, Rx all, CZ even] from "
                        "|0...0>, then <X>,<Y>,<Z> per site and <XX>,<YY>,<ZZ> per "
                        "neighbour pair; ComplexF64",
            "seconds": read_two_columns(os.path.join(
                BENCH, "benchmark_expectation_values", "timing_smoq.txt")),
        },
        "mcwf_dephasing": {
            "source": "scripts_quantum_simulations/demo_lidbladian_dm_vs_mcwf_exact_vs_trotter/"
                      "output.log",
            "workload": "XXZ chain J=(1,1,0.5), h_x=1, sigma_z dephasing gamma=0.1, "
                        "dt=0.05 x 200 first-order Trotter steps, 1000 trajectories from "
                        "|0...0>, 6 mean observables every step; ComplexF64",
            "seconds": demo_timings(os.path.join(
                REF, "scripts_quantum_simulations",
                "demo_lidbladian_dm_vs_mcwf_exact_vs_trotter", "output.log"), "MCWF+Trotter"),
        },
        "dm_dephasing_trotter": {
            "source": "scripts_quantum_simulations/demo_lidbladian_dm_vs_mcwf_exact_vs_trotter/"
                      "output.log",
            "workload": "the same chain and dissipator as a density matrix: 200 steps of "
                        "lindblad_dm_step_trotter! from |0...0><0...0|, 6 mean observables "
                        "every step; ComplexF64",
            "seconds": demo_timings(os.path.join(
                REF, "scripts_quantum_simulations",
                "demo_lidbladian_dm_vs_mcwf_exact_vs_trotter", "output.log"), "DM+Trotter"),
        },
        "dm_dephasing_exact": {
            "source": "scripts_quantum_simulations/demo_lidbladian_dm_vs_mcwf_exact_vs_trotter/"
                      "output.log",
            "workload": "200 steps of lindblad_dm_step! (dense U = exp(-iH dt), then the "
                        "dissipator), 6 mean observables every step; ComplexF64",
            "seconds": demo_timings(os.path.join(
                REF, "scripts_quantum_simulations",
                "demo_lidbladian_dm_vs_mcwf_exact_vs_trotter", "output.log"), "DM+Exact"),
        },
    }
    with open(os.path.join(HERE, "reference_timings.json"), "w") as f:
        json.dump(timings, f, indent=1)
    print(json.dumps({k: v["seconds"] for k, v in timings.items()}, indent=1))


if __name__ == "__main__":
    main()
