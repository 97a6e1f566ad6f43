"""Penalty stiffness strategies (host scalars, once per solve).

Same names, arguments and results as the reference's PenaltyStrategy
implementations (contact/penalty/strategy.py); k_pen then enters the device
assembly as a scalar argument.

  ConstantPenalty                 :34-57    k_pen at every step
  AutoBeamEIPenalty               :59-109   scale * 12 E I / L^3 / max(1, f(n_pairs)), f = n or sqrt(n)
  AutoEALPenalty                  :112-137  scale * E A / L
  ContinuationPenalty             :140-182  geometric or linear ramp to the target over ramp_steps
  DynamicPenaltyEstimateProcess   :185-262  scale * c0 * m_ii, c0 = 1 / (beta dt^2), m_ii = rho A L / 2
  _create_penalty_strategy        :265-286  AutoBeamEIPenalty if E, I, L > 0, else ConstantPenalty
"""

from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class PenaltyInput:
    step: int
    total_steps: int


@dataclass(frozen=True)
class PenaltyOutput:
    k_pen: float


class ConstantPenalty:
    """k_pen fixed for every step (strategy.py:34-57)."""

    def __init__(self, k_pen: float = 1.0) -> None:
        self._k_pen = k_pen

    def compute_k_pen(self, step: int, total_steps: int) -> float:
        return self._k_pen

    def process(self, input_data: PenaltyInput) -> PenaltyOutput:
        return PenaltyOutput(k_pen=self._k_pen)


class AutoBeamEIPenalty:
    """Bending-stiffness estimate scale * 12 E I / L^3, divided by the number of
    contact pairs (linear) or its square root (strategy.py:59-109)."""

    def __init__(self, beam_E: float, beam_I: float, L_elem: float, *, n_contact_pairs: int = 1,
                 scale: float = 0.1, scaling: str = "linear") -> None:
        if L_elem <= 0.0:
            raise ValueError(f"L_elem must be positive: {L_elem}")
        k_bend = 12.0 * beam_E * beam_I / (L_elem ** 3)
        n = max(1, n_contact_pairs)
        div = max(1.0, math.sqrt(n)) if scaling == "sqrt" else max(1.0, float(n))
        self._k_pen = scale * k_bend / div

    def compute_k_pen(self, step: int, total_steps: int) -> float:
        return self._k_pen

    def process(self, input_data: PenaltyInput) -> PenaltyOutput:
        return PenaltyOutput(k_pen=self._k_pen)


class AutoEALPenalty:
    """Axial-stiffness estimate scale * E A / L (strategy.py:112-137)."""

    def __init__(self, beam_E: float, beam_A: float, L_elem: float, *, scale: float = 1.0) -> None:
        if L_elem <= 0.0:
            raise ValueError(f"L_elem must be positive: {L_elem}")
        self._k_pen = scale * beam_E * beam_A / L_elem

    def compute_k_pen(self, step: int, total_steps: int) -> float:
        return self._k_pen

    def process(self, input_data: PenaltyInput) -> PenaltyOutput:
        return PenaltyOutput(k_pen=self._k_pen)


class ContinuationPenalty:
    """Penalty ramped to its target over ramp_steps steps (strategy.py:140-182):
    geometric k = min(target * start * ratio^step, target), ratio =
    (1/start)^(1/ramp_steps); linear k = target (start + (1 - start) step / ramp_steps)."""

    def __init__(self, k_pen_target: float, *, start_fraction: float = 0.01, ramp_steps: int = 5,
                 mode: str = "geometric") -> None:
        self._target = k_pen_target
        self._start = max(start_fraction, 1e-30)
        self._ramp = max(1, ramp_steps)
        self._mode = mode

    def compute_k_pen(self, step: int, total_steps: int) -> float:
        if step >= self._ramp:
            return self._target
        if self._mode == "geometric":
            ratio = (1.0 / self._start) ** (1.0 / self._ramp)
            return min(self._target * self._start * ratio ** step, self._target)
        return self._target * (self._start + (1.0 - self._start) * step / self._ramp)

    def process(self, input_data: PenaltyInput) -> PenaltyOutput:
        return PenaltyOutput(k_pen=self.compute_k_pen(input_data.step, input_data.total_steps))


@dataclass(frozen=True)
class DynamicPenaltyEstimateInput:
    rho_inf: float
    dt: float
    rho: float
    A: float
    L_elem: float
    scale: float = 0.5


@dataclass(frozen=True)
class DynamicPenaltyEstimateOutput:
    k_pen: float
    c0: float
    m_ii: float
    c0_m_ii: float


class DynamicPenaltyEstimateProcess:
    """k_pen = scale * c0 * m_ii from the Generalized-alpha effective mass
    stiffness (Chung-Hulbert parameters; strategy.py:185-262)."""

    def process(self, input_data: DynamicPenaltyEstimateInput) -> DynamicPenaltyEstimateOutput:
        rho_inf = input_data.rho_inf
        alpha_m = (2.0 * rho_inf - 1.0) / (rho_inf + 1.0)
        alpha_f = rho_inf / (rho_inf + 1.0)
        beta = 0.25 * (1.0 - alpha_m + alpha_f) ** 2
        c0 = 1.0 / (beta * input_data.dt ** 2)
        m_ii = input_data.rho * input_data.A * input_data.L_elem / 2.0
        c0_m_ii = c0 * m_ii
        return DynamicPenaltyEstimateOutput(k_pen=input_data.scale * c0_m_ii, c0=c0, m_ii=m_ii, c0_m_ii=c0_m_ii)


def _create_penalty_strategy(*, k_pen: float = 1.0, beam_E: float = 0.0, beam_I: float = 0.0,
                             beam_L: float = 0.0):
    """AutoBeamEIPenalty when E, I and L are all positive, else ConstantPenalty(k_pen)
    (strategy.py:265-286)."""
    if beam_E > 0.0 and beam_I > 0.0 and beam_L > 0.0:
        return AutoBeamEIPenalty(beam_E=beam_E, beam_I=beam_I, L_elem=beam_L)
    return ConstantPenalty(k_pen=k_pen)
