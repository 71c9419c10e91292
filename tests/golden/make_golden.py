"""Write the golden fixtures (tests/golden/*.npz) from the oracle.

    python tests/golden/make_golden.py [case ...]

Each fixture holds the oracle's state after the case's steps (see cases.py).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (HERE, os.path.join(REPO, "dimcv-sph_amd"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

import cases  # noqa: E402


def main(names):
    for name in names or list(cases.CASES):
        out = cases.run_oracle(name)
        np.savez_compressed(cases.fixture_path(name), **out)
        size = os.path.getsize(cases.fixture_path(name))
        print(f"{name}: particle_num={int(out['particle_num'])} samples={int(out['num_samples'])} "
              f"kv={int(out['kv'])} kp={int(out['kp'])} -> {size} bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
