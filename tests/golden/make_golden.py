"""Writes the golden fixtures in this folder (run from the repository root:
`python tests/golden/make_golden.py`). Everything is produced from this
repository's own code, nothing is imported from the reference:

* known_answers.json -- the known-answer data the reference's tests and
  tutorials hold (tutorials/unittest_robast.py:201-205,415-426,527-561,
  tutorials/EdmundOptics.C:57-73, misc/schottzemax-20180601.agf:743-746)
  and the Random123 Philox4x32-10 vectors;
* dc_facets.json -- the 85 Davies-Cotton facet centres and placement
  matrices (tutorials/DaviesCotton.C:75-124 construction);
* c3_coefficients.json -- the seeded Schwarzschild-Couder conics and
  polynomial coefficients;
* c1_sample.npz, c2_sample.npz, c3_sample.npz, c4_sample.npz -- input rays
  and the FP64 oracle's per-ray results (status, point count, last node,
  terminal point, time, direction). For C1, 1 % of the focused rays are
  first checked against the exact 50-digit reflection chain, so the
  fixture is not only the oracle's word.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import c1_exact_chain, copy_soa  # noqa: E402
from oracle import oracle as O  # noqa: E402
from robast_amd import _abi as A  # noqa: E402
from robast_amd import configs  # noqa: E402

FIELDS = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node")

KNOWN = {
    "philox4x32_10": [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
        {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
         "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
        {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
         "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
    ],
    "tgraph": [
        {"points_nm": [[400, 1.6], [500, 1.5]], "x_nm": 450, "y": 1.55, "exact": True},
        {"points_nm": [[300, 0.0], [500, 0.5]], "x_nm": 400, "y": 0.25, "exact": False},
    ],
    "nbk7": {
        "sellmeier": [1.03961212, 0.231792344, 1.01046945, 0.00600069867, 0.0200179144, 103.560653],
        "table_nm": [[2325.4, 1.489210], [1970.1, 1.494950], [1529.6, 1.500910], [1060.0, 1.506690],
                     [1014.0, 1.507310], [852.1, 1.509800], [706.5, 1.512890], [656.3, 1.514320],
                     [643.8, 1.514720], [632.8, 1.515090], [589.3, 1.516730], [587.6, 1.516800],
                     [546.1, 1.518720], [486.1, 1.522380], [480.0, 1.522830], [435.8, 1.526680],
                     [404.7, 1.530240], [365.0, 1.536270], [334.1, 1.542720], [312.6, 1.548620]],
        "table_decimals": 4,
        "abbe": 64.17,
        "nd_agf": 1.51680,
    },
    "ohara": [
        {"name": "S-FSL5", "sellmeier": [1.17447043e0, 1.40056154e-2, 1.19272435e0, 8.41855181e-3,
                                         -5.81790767e-2, 1.29599726e2], "lambda_um": 0.58756, "n": 1.48749},
        {"name": "S-TIH13", "sellmeier": [1.62224674e0, 2.93844589e-1, 1.99225164e0, 1.18368386e-2,
                                          5.90208025e-2, 1.71959976e2], "lambda_um": 0.58756, "n": 1.74077},
    ],
}


def sample_c1():
    mgr = configs.c1_scene()
    src = configs.c1_rays(1.0, nside=64).soa()
    return mgr, src


def sample_c2():
    mgr = configs.c2_scene()
    return mgr, O.generate(configs.c2_raygen(), 4000)


def sample_c3():
    mgr = configs.c3_scene()
    return mgr, O.generate(configs.c3_raygen(), 2000)


def sample_c4():
    mgr = configs.c4_scene()
    return mgr, O.generate(configs.c4_raygen(mgr.c4_dz), 2000)


SAMPLES = {"c1": sample_c1, "c2": sample_c2, "c3": sample_c3, "c4": sample_c4}


def main():
    with open(os.path.join(HERE, "known_answers.json"), "w") as fh:
        json.dump(KNOWN, fh, indent=1)
    facets = []
    for i, (x, y) in enumerate(configs.dc_facet_positions()):
        mat = configs._dc_facet_matrix(x, y)
        facets.append({"x": x, "y": y, "rot": [float(v) for v in np.asarray(mat.rot).ravel()],
                       "tr": [float(v) for v in mat.tr]})
    with open(os.path.join(HERE, "dc_facets.json"), "w") as fh:
        json.dump(facets, fh, indent=0)
    coef = {k: {"conic": float(conic), "poly": [float(v) for v in ks]}
            for k, (conic, ks) in configs.c3_coefficients().items()}
    with open(os.path.join(HERE, "c3_coefficients.json"), "w") as fh:
        json.dump(coef, fh, indent=1)
    for name, make in SAMPLES.items():
        mgr, src = make()
        out = copy_soa(src)
        O.trace(mgr.flat(), out, mgr.trace_opts(), nthreads=8)
        if name == "c1":
            idx = np.nonzero(out["status"] == A.RB_FOCUS)[0][::100]
            for i in idx:
                end, tt = c1_exact_chain(mgr.flat(), src, i, 2e-6)
                for k, key in enumerate(("x", "y", "z")):
                    assert abs(float(end[k]) - out[key][i]) < 1e-9
                assert abs(float(tt) - out["t"][i]) < 1e-18
            print(f"c1: {len(idx)} focused rays checked against the exact chain")
        np.savez_compressed(os.path.join(HERE, f"{name}_sample.npz"),
                            **{f"in_{k}": src[k] for k in FIELDS}, **{f"out_{k}": out[k] for k in FIELDS})
        print(name, len(src["x"]), np.bincount(out["status"], minlength=6).tolist())


if __name__ == "__main__":
    main()
