"""Penalty strategies (xkep-cae_amd/penalty.py) against the known answers of
the reference's own tests (contact/penalty/tests/test_strategy.py,
tests/contact/test_dynamic_penalty_estimate.py), and the squeeze workload's
penalty against the dynamic estimate."""

import math

import pytest

from conftest import pkg


def test_auto_beam_ei_penalty_known_answers():
    P = pkg("penalty")
    k_bend = 12.0 * 200e9 * 1e-12 / 0.01 ** 3
    assert P.AutoBeamEIPenalty(beam_E=200e9, beam_I=1e-12, L_elem=0.01).compute_k_pen(0, 10) == pytest.approx(
        0.1 * k_bend)
    assert P.AutoBeamEIPenalty(beam_E=200e9, beam_I=1e-12, L_elem=0.01, n_contact_pairs=6).compute_k_pen(
        0, 10) == pytest.approx(0.1 * k_bend / 6)
    assert P.AutoBeamEIPenalty(beam_E=200e9, beam_I=1e-12, L_elem=0.01, n_contact_pairs=9,
                               scaling="sqrt").compute_k_pen(0, 10) == pytest.approx(0.1 * k_bend / 3.0)
    p = P.AutoBeamEIPenalty(beam_E=200e9, beam_I=1e-12, L_elem=0.01)
    assert p.compute_k_pen(0, 10) == p.compute_k_pen(5, 10)
    assert isinstance(p.process(P.PenaltyInput(step=0, total_steps=10)), P.PenaltyOutput)
    with pytest.raises(ValueError):
        P.AutoBeamEIPenalty(beam_E=200e9, beam_I=1e-12, L_elem=0.0)


def test_axial_constant_and_factory():
    P = pkg("penalty")
    assert P.AutoEALPenalty(beam_E=200e9, beam_A=1e-6, L_elem=0.01, scale=1.0).compute_k_pen(0, 1) == pytest.approx(
        200e9 * 1e-6 / 0.01)
    c = P.ConstantPenalty(k_pen=42.0)
    assert c.compute_k_pen(0, 10) == 42.0 and c.compute_k_pen(5, 10) == 42.0
    assert P.ConstantPenalty(k_pen=1e6).process(P.PenaltyInput(0, 1)).k_pen == 1e6
    s = P._create_penalty_strategy(k_pen=1e5)
    assert isinstance(s, P.ConstantPenalty) and s.compute_k_pen(0, 1) == 1e5
    assert isinstance(P._create_penalty_strategy(beam_E=200e9, beam_I=1e-12, beam_L=0.01), P.AutoBeamEIPenalty)
    assert isinstance(P._create_penalty_strategy(k_pen=1e5, beam_E=200e9, beam_I=1e-12), P.ConstantPenalty)


def test_continuation_penalty_ramps():
    P = pkg("penalty")
    g = P.ContinuationPenalty(k_pen_target=1e5, start_fraction=0.01, ramp_steps=5, mode="geometric")
    assert g.compute_k_pen(0, 10) == pytest.approx(0.01 * 1e5)
    assert g.compute_k_pen(5, 10) == pytest.approx(1e5)
    vals = [g.compute_k_pen(i, 10) for i in range(8)]
    assert all(b >= a for a, b in zip(vals, vals[1:]))
    lin = P.ContinuationPenalty(k_pen_target=1e5, start_fraction=0.1, ramp_steps=10, mode="linear")
    assert lin.compute_k_pen(0, 10) == pytest.approx(0.1 * 1e5)
    assert lin.compute_k_pen(5, 10) == pytest.approx(1e5 * (0.1 + 0.9 * 5 / 10))
    assert lin.compute_k_pen(10, 10) == pytest.approx(1e5)
    assert lin.compute_k_pen(100, 100) == pytest.approx(1e5)


def three_point_bend_estimate(scale=0.5):
    """The reference test's three-point-bend conditions: L 100, d 2, E 200e3,
    rho 7.85e-9, dt = T1/40, 20 elements, rho_inf 0.9."""
    P = pkg("penalty")
    E, d, L, rho = 200e3, 2.0, 100.0, 7.85e-9
    Iy = math.pi * d ** 4 / 64
    A = math.pi * d ** 2 / 4
    f1 = (math.pi / (2.0 * L ** 2)) * math.sqrt(E * Iy / (rho * A))
    dt = 1.0 / f1 / 40.0
    out = P.DynamicPenaltyEstimateProcess().process(
        P.DynamicPenaltyEstimateInput(rho_inf=0.9, dt=dt, rho=rho, A=A, L_elem=L / 20, scale=scale))
    return out, 48.0 * E * Iy / L ** 3


def test_dynamic_penalty_estimate():
    P = pkg("penalty")
    out, k_static = three_point_bend_estimate()
    assert 10.0 < out.k_pen < 100.0 and 30.0 < out.c0_m_ii < 100.0
    assert out.k_pen / k_static > 2.0
    assert out.c0_m_ii == pytest.approx(out.c0 * out.m_ii, rel=1e-15)
    assert three_point_bend_estimate(1.0)[0].k_pen == pytest.approx(2.0 * out.k_pen, rel=1e-15)
    # positive definiteness margin: k_pen < (1 - alpha_m) c0 m_ii at the default scale
    rho_inf = 0.9
    am = (2.0 * rho_inf - 1.0) / (rho_inf + 1.0)
    o = P.DynamicPenaltyEstimateProcess().process(
        P.DynamicPenaltyEstimateInput(rho_inf=rho_inf, dt=1e-5, rho=7.85e-9, A=math.pi, L_elem=5.0))
    assert o.k_pen < (1.0 - am) * o.c0_m_ii


def test_squeeze_workload_penalty_is_the_dynamic_estimate():
    P, W = pkg("penalty"), pkg("workloads")
    sq = W.squeeze_case(100.0)
    A = math.pi * 1.0 ** 2
    dt = sq["t_total"] * sq["cfg"]["dt_max_fraction"]
    out = P.DynamicPenaltyEstimateProcess().process(
        P.DynamicPenaltyEstimateInput(rho_inf=sq["rho_inf"], dt=dt, rho=W.RHO_STEEL, A=A, L_elem=20.0 / 20))
    assert sq["k_pen"] == pytest.approx(out.k_pen, rel=1e-14)
