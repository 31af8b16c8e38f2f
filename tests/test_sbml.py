"""The SBML-subset reader (amici_amd/sbml.py) against the reaction networks it
must reproduce: AMICI's stoichiometry rule S[species, reaction] += sign *
stoichiometry (python/sdist/amici/importers/sbml/__init__.py:1362-1428) and
mass-action rate laws."""
from __future__ import annotations

import math

import numpy as np
import pytest

from amici_amd import codegen, networks, sbml

DATA = __import__("pathlib").Path(__file__).resolve().parents[1] / "amici_amd" / "data"


def _same(a, b):
    np.testing.assert_array_equal(a.stoichiometry(), b.stoichiometry())
    assert [r.rate for r in a.reactions] == [r.rate for r in b.reactions]
    np.testing.assert_allclose(a.theta_nominal, b.theta_nominal, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(a.x0_const, b.x0_const)
    np.testing.assert_array_equal(a.x0_fixed, b.x0_fixed)
    np.testing.assert_allclose(a.observable_matrix, b.observable_matrix)
    assert a.noise == b.noise
    assert codegen.model_tables(a) == codegen.model_tables(b)


def test_robertson_written_as_sbml():
    net = sbml.read_sbml(DATA / "robertson.xml", observables={"yA": "A", "yC": "C"})
    _same(net, networks.robertson())
    assert net.species == ["A", "B", "C"]
    assert net.parameter_names == ["k1", "k2", "r3_k3"]  # local parameter renamed


@pytest.mark.parametrize("make", [networks.robertson, networks.network_n40,
                                  lambda: networks.linear_chain(4)])
def test_round_trip(make):
    net = make()
    text = sbml.write_sbml(net)
    back = sbml.read_sbml(text, fixed_parameters=net.fixed_parameter_names,
                          noise="lognormal" if net.noise == networks.NOISE_LOGNORMAL else "normal")
    _same(back, net)
    np.testing.assert_allclose(back.k_nominal, net.k_nominal)


MODEL = """<?xml version="1.0"?>
<sbml xmlns="http://www.sbml.org/sbml/level3/version1/core" level="3" version="1">
 <model id="m">
  <listOfCompartments><compartment id="c" size="{size}"/></listOfCompartments>
  <listOfSpecies>
   <species id="A" compartment="c" initialConcentration="1"/>
   <species id="B" compartment="c" initialConcentration="0"/>
  </listOfSpecies>
  <listOfParameters><parameter id="kf" value="2"/><parameter id="kr" value="0.5"/>
   <parameter id="Km" value="1"/></listOfParameters>
  <listOfReactions>
   <reaction id="r" reversible="true">
    <listOfReactants><speciesReference species="A"/></listOfReactants>
    <listOfProducts><speciesReference species="B"/></listOfProducts>
    <kineticLaw><math xmlns="http://www.w3.org/1998/Math/MathML">{law}</math></kineticLaw>
   </reaction>
  </listOfReactions>
 </model>
</sbml>"""

REV = ("<apply><minus/><apply><times/><ci>kf</ci><ci>A</ci></apply>"
       "<apply><times/><ci>kr</ci><ci>B</ci></apply></apply>")
MM = ("<apply><divide/><apply><times/><ci>kf</ci><ci>A</ci></apply>"
      "<apply><plus/><ci>Km</ci><ci>A</ci></apply></apply>")


def test_reversible_law_splits_into_two_reactions():
    net = sbml.read_sbml(MODEL.format(size=1, law=REV), observables={"y": "A + 2*B"})
    assert [(r.reactants, r.products) for r in net.reactions] == [((0,), (1,)), ((1,), (0,))]
    np.testing.assert_allclose(10 ** net.theta_nominal, [2.0, 0.5])
    np.testing.assert_array_equal(net.observable_matrix, [[1.0, 2.0]])


@pytest.mark.parametrize("size,law,obs", [(1, MM, {"y": "A"}), (2, REV, {"y": "A"}),
                                          (1, REV, {"y": "A*B"})])
def test_outside_the_subset_is_rejected(size, law, obs):
    with pytest.raises(sbml.SbmlError):
        sbml.read_sbml(MODEL.format(size=size, law=law), observables=obs)


def test_noise_distribution_names():
    """AMICI's continuous noise-distribution names select the noise model;
    others are refused."""
    for name, want in (("normal", networks.NOISE_NORMAL), ("log-normal", networks.NOISE_LOGNORMAL),
                       ("lognormal", networks.NOISE_LOGNORMAL),
                       ("log10-normal", networks.NOISE_LOG10NORMAL),
                       ("laplace", networks.NOISE_LAPLACE), ("log-laplace", networks.NOISE_LOGLAPLACE),
                       ("log10-laplace", networks.NOISE_LOG10LAPLACE)):
        net = sbml.read_sbml(DATA / "robertson.xml", observables={"yA": "A", "yC": "C"}, noise=name)
        assert net.noise == want
    with pytest.raises(sbml.SbmlError):
        sbml.read_sbml(DATA / "robertson.xml", observables={"yA": "A"}, noise="binomial")


def test_log10_residuals():
    """Residuals on the log10 scale for the log10 noise models (fres,
    rdata.cpp:985-996) and their sensitivities (fsres)."""
    import torch
    from amici_amd import rdata
    y = torch.tensor([[[2.0, 0.5]]], dtype=torch.float64)
    m = torch.tensor([[[3.0, float("nan")]]], dtype=torch.float64)
    s = torch.full_like(y, 0.2)
    for noise in (networks.NOISE_LOG10NORMAL, networks.NOISE_LOG10LAPLACE):
        r = rdata.residuals(y, m, s, noise)
        assert r[0, 0, 0].item() == pytest.approx((math.log10(3.0) - math.log10(2.0)) / 0.2)
        assert r[0, 0, 1].item() == 0.0
        sy = torch.tensor([[[[1.0, 1.0]]]], dtype=torch.float64)  # [nb, nt, np, ny]
        sr = rdata.residual_sensitivities(y, sy, m, s, noise)
        assert sr[0, 0, 0, 0].item() == pytest.approx(-1.0 / (2.0 * math.log(10.0) * 0.2))
