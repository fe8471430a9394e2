"""Writes the regression fixtures of tests/golden/ from the CPU oracle.

These vectors pin the oracle against drift between runs and machines (they are produced by
the oracle, so they are regression pins; correctness pins are the closed-form, invariant,
brute-force and independent-formulation checks in tests/test_oracle.py). The reference has
no tests or fixtures of its own for this path.
Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from minotert_amd import configs as cfg  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def main():
    sums = {}
    for cid in (0, 1):
        v, m, a = orc.generate_scene(cfg.scene_desc(cid))
        sums[str(cid)] = hashlib.sha256(v.tobytes() + m.tobytes() + a.tobytes()).hexdigest()
    with open(os.path.join(HERE, "scene_checksums.json"), "w") as f:
        json.dump(sums, f, indent=1)
    pcg_in = np.array([0, 1, 2, 3, 0xFFFFFFFF, 0x80000000, 12345, 0x9E3779B9, 0x5EED0000, 0xDEADBEEF], np.uint32)
    pcg_out = np.array([orc.pcg(int(x)) for x in pcg_in], np.uint32)
    rays32 = orc.primary_rays(cfg.camera(0, width=32, height=32))
    v, m, a = orc.generate_scene(cfg.scene_desc(0))
    probe = np.concatenate([rays32[::8], orc.random_rays(99, 10.0, 128)])
    hits = orc.closest_brute(v, probe)
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), pcg_in=pcg_in, pcg_out=pcg_out,
                        primary_rays_32=rays32, probe_rays=probe,
                        probe_hits=np.ascontiguousarray(hits).view(np.uint32).reshape(-1, 4))
    print("wrote", sums)


if __name__ == "__main__":
    main()
