"""Regenerate the golden fixtures in this directory (CPU only).

    python tests/golden/make_golden.py

Writes, for three of the benchmark circuits (seeded generators in
clifft_amd.workloads):
  <name>.stim      the circuit text
  <name>.prog      the compiled Program blob (Program.to_bytes)
  <name>.npz       records / expectation values / gradients computed by the
                   CPU oracle on that blob with fixed seeds
The fixtures pin the compiler's output and the oracle's results against
regressions; the device executor is checked against the same .npz files.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

SHOTS = {"clifford_t20": 256, "surface_d3": 2048}
SEED = 42


def circuits():
    from clifft_amd import workloads as W
    build, coeffs, _ = W.rz_exp_val()
    thetas = np.random.default_rng(SEED).uniform(0, 2 * np.pi, (2, 16))
    return {
        "clifford_t20": (W.random_clifford_t(), {}),
        "surface_d3": (W.surface_code_d3(), {}),
        "rz24": (build(thetas[0]), {"keep_rotations": True, "thetas": thetas, "coeffs": coeffs}),
    }


def main():
    from clifft_amd.compiler import compile as cc
    from gradprog import oracle_exp_vals, oracle_param_shift
    from oracle import oracle as orc
    for name, (text, opts) in circuits().items():
        kw = {"keep_rotations": True} if opts.get("keep_rotations") else {}
        prog = cc(text, **kw)
        with open(os.path.join(HERE, name + ".stim"), "w") as f:
            f.write(text + "\n")
        with open(os.path.join(HERE, name + ".prog"), "wb") as f:
            f.write(prog.to_bytes())
        if name in SHOTS:
            m, d, o, e, _ = orc.sample(prog, SHOTS[name], seed=SEED, per_shot=True)
            first = orc.sample(prog, 1, seed=SEED, per_shot=False)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), seed=SEED,
                                measurements=np.packbits(m, axis=1, bitorder="little"),
                                detectors=np.packbits(d, axis=1, bitorder="little"),
                                observables=np.packbits(o, axis=1, bitorder="little"),
                                num_measurements=prog.num_measurements, num_detectors=prog.num_detectors,
                                num_observables=prog.num_observables,
                                single_stream_first=np.concatenate([first[0][0], first[1][0], first[2][0]]))
        else:
            th, c = opts["thetas"], opts["coeffs"]
            ev = np.stack([oracle_exp_vals(orc, prog, t) for t in th])
            gr = np.stack([oracle_param_shift(orc, prog, t, c) for t in th])
            np.savez_compressed(os.path.join(HERE, name + ".npz"), thetas=th, coeffs=c, exp_vals=ev, grads=gr)
        print(name, prog.peak_rank, len(prog.bytecode), "instructions")


if __name__ == "__main__":
    main()
