"""Writes the committed golden fixtures of this directory.

gaussian_seed152.json: the first 64 raw MT19937 outputs and the first 64
N(0, 1) sketch draws for seed 152 (the sketch seed of
FactorManager.cpp:58-71), produced by the oracle's C restatement of
MERSENNETWISTER.h (randInt / randNorm, :170-194).  Tests compare the
device-side generator and a numpy restatement of the Box-Muller step with it.

relaxed_supernodes_handworked.json is written by hand (its derivations are
in tests/test_relaxed_handworked.py), not by this script.

Run from the repository root: python tests/golden/make_goldens.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def main():
    from oracle.oracle import gaussian, mt_output

    raw = [int(mt_output(152, k + 1)) for k in range(64)]  # 1-based index
    g = gaussian(152, 64).tolist()
    with open(os.path.join(HERE, "gaussian_seed152.json"), "w") as f:
        json.dump({"seed": 152, "mt_outputs": raw, "normals": g}, f, indent=1)



if __name__ == "__main__":
    main()
