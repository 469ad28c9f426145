"""Golden fixtures (tests/golden/, written by make_golden.py): the
known-answer data of the reference's tests and tutorials, the Davies-Cotton
facet table, the seeded Schwarzschild-Couder coefficients, and per-ray
results for samples of the four tracing configurations.

CPU: the oracle and the configuration builders reproduce the fixtures
exactly. GPU: librbgpu, through the C ABI, reproduces the fixtures' per-ray
results (status, point count, last node identical; terminal point within
1e-7 cm, direction within 1e-9)."""
import json
import math
import os

import numpy as np
import pytest

from helpers import FIELDS, compare, copy_soa
from oracle import oracle as O
from robast_amd import configs
from robast_amd import geometry as G
from robast_amd.flatten import flatten
from robast_amd.materials import ARefractiveIndex, ASellmeierFormula, TGraph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _json(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def _sample(name):
    z = np.load(os.path.join(GOLDEN, f"{name}_sample.npz"))
    return {k: z[f"in_{k}"].copy() for k in FIELDS}, {k: z[f"out_{k}"].copy() for k in FIELDS}


def _scene(name):
    return {"c1": configs.c1_scene, "c2": configs.c2_scene, "c3": configs.c3_scene, "c4": configs.c4_scene}[name]()


def _inputs(name, mgr, n):
    if name == "c1":
        return configs.c1_rays(1.0, nside=64).soa()
    gen = {"c2": configs.c2_raygen, "c3": configs.c3_raygen}.get(name)
    return O.generate(gen() if gen else configs.c4_raygen(mgr.c4_dz), n)


def _lens(index):
    top = G.AOpticalComponent("world", G.TGeoBBox("w", 1, 1, 1))
    lens = G.ALens("lens", G.TGeoBBox("l", 0.5, 0.5, 0.5))
    lens.SetRefractiveIndex(index)
    top.AddNode(lens, 1)
    return flatten(top)


def test_known_answers():
    ka = _json("known_answers.json")
    for v in ka["philox4x32_10"]:
        assert list(O.philox(v["ctr"], v["key"])) == v["out"]
    nm = G.nm
    for v in ka["tgraph"]:
        g = TGraph()
        for k, (x, y) in enumerate(v["points_nm"]):
            g.SetPoint(k, x * nm, y)
        idx = ARefractiveIndex()
        idx.SetRefractiveIndex(g)
        got = O.lens_index(_lens(idx), 1, v["x_nm"] * nm)
        assert got == v["y"] if v["exact"] else abs(got - v["y"]) < 1e-12
    b = ka["nbk7"]
    s = ASellmeierFormula(*b["sellmeier"])
    flat = _lens(s)
    for lam, n in b["table_nm"]:
        assert round(O.lens_index(flat, 1, lam * nm) - n, b["table_decimals"]) == 0
    assert round(s.GetAbbeNumber() - b["abbe"], 1) == 0
    assert round(O.lens_index(flat, 1, 589.3 * nm) / b["nd_agf"] - 1, 4) == 0
    for v in ka["ohara"]:
        assert round(O.lens_index(_lens(ASellmeierFormula(*v["sellmeier"])), 1, v["lambda_um"] * G.um), 5) == v["n"]


def test_dc_facet_table():
    facets = _json("dc_facets.json")
    pos = configs.dc_facet_positions()
    assert len(facets) == len(pos) == configs.C2_NFACETS
    for f, (x, y) in zip(facets, pos):
        assert (f["x"], f["y"]) == (x, y)
        mat = configs._dc_facet_matrix(x, y)
        assert np.array_equal(np.asarray(mat.rot).ravel(), f["rot"])
        assert np.array_equal(np.asarray(mat.tr), f["tr"])
        # each facet axis points at (0, 0, 2f)
        axis = np.asarray(f["rot"]).reshape(3, 3)[:, 2]
        to = np.array([0.0, 0.0, 2 * configs.C2_F]) - np.asarray(f["tr"])
        assert np.allclose(axis, to / np.linalg.norm(to), atol=1e-12)


def test_c3_coefficients():
    coef = _json("c3_coefficients.json")
    for name, (conic, ks) in configs.c3_coefficients().items():
        assert coef[name]["conic"] == conic
        assert coef[name]["poly"] == list(ks)
        assert -1.0 <= conic <= 0.0


@pytest.mark.parametrize("name", ["c1", "c2", "c3", "c4"])
def test_oracle_reproduces_fixture(name):
    src, want = _sample(name)
    mgr = _scene(name)
    regen = _inputs(name, mgr, len(src["x"]))
    for k in FIELDS:
        assert np.array_equal(regen[k], src[k]), k
    got = copy_soa(src)
    O.trace(mgr.flat(), got, mgr.trace_opts(), nthreads=8)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1", "c2", "c3", "c4"])
def test_gpu_reproduces_fixture(ctx, name):
    from robast_amd._lib import make_rays
    print(f"golden {name}", flush=True)
    src, want = _sample(name)
    mgr = _scene(name)
    got = copy_soa(src)
    ctx.trace(ctx.upload(mgr.flat()), make_rays(got), len(got["x"]), mgr.trace_opts())
    compare(got, want)
    assert math.isclose(got["t"].sum(), want["t"].sum(), rel_tol=1e-12)
