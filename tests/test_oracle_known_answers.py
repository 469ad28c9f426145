"""Pins the CPU oracle against known answers: Random123 Philox vectors, the
reference's own unit-test values (tutorials/unittest_robast.py) and
manufacturer index data quoted in the reference (tutorials/EdmundOptics.C,
misc/schottzemax-20180601.agf)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd.flatten import flatten
from robast_amd.materials import ARefractiveIndex, ASchottFormula, ASellmeierFormula, TGraph

nm, um, mm, m = G.nm, G.um, G.mm, G.m


def test_units_match_reference():
    # include/AOpticsManager.h:59-71 (cm base unit)
    assert G.m == 100.0 and G.cm == 1.0
    assert G.mm == 1e-3 * 100.0 and G.um == 1e-6 * 100.0 and G.nm == 1e-9 * 100.0


@pytest.mark.parametrize("ctr,key,expect", [
    # Random123 known-answer vectors for philox4x32-10
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert list(O.philox(ctr, key)) == expect


def test_uniform_range_and_moments():
    u = np.array([O.uniform(7, i, k) for i in range(2000) for k in range(5)])
    assert u.min() > 0 and u.max() <= 1
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / len(u))
    assert abs(u.var() - 1 / 12) < 0.01


def _lens_scene(index):
    top = G.AOpticalComponent("world", G.TGeoBBox("w", 1, 1, 1))
    lens = G.ALens("lens", G.TGeoBBox("l", 0.5, 0.5, 0.5))
    lens.SetRefractiveIndex(index)
    top.AddNode(lens, 1)
    return flatten(top)


def test_tgraph_eval_exact_reference_values():
    # testRefractiveIndex (unittest_robast.py:415-426): n(450 nm) == 1.55 exactly
    g = TGraph()
    g.SetPoint(0, 400 * nm, 1.6)
    g.SetPoint(1, 500 * nm, 1.5)
    idx = ARefractiveIndex()
    idx.SetRefractiveIndex(g)
    assert idx.GetRefractiveIndex(450 * nm) == 1.55
    flat = _lens_scene(idx)
    assert O.lens_index(flat, 1, 450 * nm) == 1.55
    # testMirrorReflection: TGraph (300 nm, 0) (500 nm, 0.5) -> 0.25 at 400 nm
    g2 = TGraph()
    g2.SetPoint(0, 300 * nm, 0.0)
    g2.SetPoint(1, 500 * nm, 0.5)
    assert abs(g2.Eval(400 * nm) - 0.25) < 1e-12


def test_tgraph_eval_oracle_matches_host_including_extrapolation():
    rng = np.random.default_rng(1)
    xs = np.sort(rng.uniform(0, 10, 7))
    ys = rng.normal(size=7)
    g = TGraph()
    for k, i in enumerate(rng.permutation(7)):  # insertion order must not matter
        g.SetPoint(k, xs[i], ys[i])
    idx = ARefractiveIndex()
    idx.SetRefractiveIndex(g)
    flat = _lens_scene(idx)
    for x in list(rng.uniform(-5, 15, 200)) + list(xs):
        assert O.lens_index(flat, 1, x) == pytest.approx(g.Eval(x), rel=0, abs=1e-12)


def test_single_point_graph_is_constant():
    idx = ARefractiveIndex(1.7)
    flat = _lens_scene(idx)
    for lam in (200 * nm, 400 * nm, 900 * nm):
        assert O.lens_index(flat, 1, lam) == 1.7


# N-BK7 from the SCHOTT catalogue (unittest_robast.py:530-561)
NBK7 = (1.03961212, 0.231792344, 1.01046945, 0.00600069867, 0.0200179144, 103.560653)
NBK7_TABLE = ((2325.4, 1.489210), (1970.1, 1.494950), (1529.6, 1.500910), (1060.0, 1.506690),
              (1014.0, 1.507310), (852.1, 1.509800), (706.5, 1.512890), (656.3, 1.514320),
              (643.8, 1.514720), (632.8, 1.515090), (589.3, 1.516730), (587.6, 1.516800),
              (546.1, 1.518720), (486.1, 1.522380), (480.0, 1.522830), (435.8, 1.526680),
              (404.7, 1.530240), (365.0, 1.536270), (334.1, 1.542720), (312.6, 1.548620))


def test_sellmeier_nbk7_table_and_abbe():
    s = ASellmeierFormula(*NBK7)
    flat = _lens_scene(s)
    for lam_nm, n in NBK7_TABLE:
        assert round(O.lens_index(flat, 1, lam_nm * nm) - n, 4) == 0
        assert O.lens_index(flat, 1, lam_nm * nm) == pytest.approx(s.GetRefractiveIndex(lam_nm * nm), abs=1e-15)
    assert round(s.GetAbbeNumber() - 64.17, 1) == 0
    # AGF catalogue nD (misc/schottzemax-20180601.agf:743-746)
    assert round(s.GetRefractiveIndex(589.3 * nm) / 1.51680 - 1, 4) == 0


def test_sellmeier_ohara_glasses():
    # tutorials/EdmundOptics.C:57-73 "should be 1.48749" / "should be 1.74077"
    fsl5 = ASellmeierFormula(1.17447043e0, 1.40056154e-2, 1.19272435e0, 8.41855181e-3, -5.81790767e-2, 1.29599726e2)
    tih13 = ASellmeierFormula(1.62224674e0, 2.93844589e-1, 1.99225164e0, 1.18368386e-2, 5.90208025e-2,
                              1.71959976e2)
    assert round(O.lens_index(_lens_scene(fsl5), 1, 0.58756 * um), 5) == 1.48749
    assert round(O.lens_index(_lens_scene(tih13), 1, 0.58756 * um), 5) == 1.74077


def test_schott_formula_matches_closed_form_and_sellmeier_limit():
    # A0 + A1 l^2 + sum A_k l^-2k : a pure constant index when A1..A5 = 0
    c = ASchottFormula(2.25, 0, 0, 0, 0, 0)
    assert O.lens_index(_lens_scene(c), 1, 500 * nm) == 1.5
    a = (2.2718929, -1.0108077e-2, 1.0592509e-2, 2.0816965e-4, -7.6472538e-6, 4.9240991e-7)
    s = ASchottFormula(*a)
    flat = _lens_scene(s)
    for lam in (350 * nm, 500 * nm, 800 * nm):
        l = lam / um
        n = math.sqrt(a[0] + a[1] * l ** 2 + a[2] * l ** -2 + a[3] * l ** -4 + a[4] * l ** -6 + a[5] * l ** -8)
        assert O.lens_index(flat, 1, lam) == pytest.approx(n, rel=1e-14)
    # this Schott-form fit of N-BK7 agrees with the Sellmeier table to 1e-4 in the visible
    assert abs(O.lens_index(flat, 1, 587.6 * nm) - 1.5168) < 2e-4


def test_absorption_length_conversion():
    # include/ARefractiveIndex.h:55-62
    lam, a = 400 * nm, 1 * mm
    k = ARefractiveIndex.AbsorptionLengthToExtinctionCoefficient(a, lam)
    assert ARefractiveIndex.ExtinctionCoefficientToAbsorptionLength(k, lam) == pytest.approx(a, rel=1e-15)
    idx = ARefractiveIndex(1.0, k)
    assert idx.GetAbsorptionLength(lam) == pytest.approx(a, rel=1e-15)
    # one-point k graph: the length scales with lambda (lambda / (4 pi k))
    assert idx.GetAbsorptionLength(2 * lam) == pytest.approx(2 * a, rel=1e-15)
