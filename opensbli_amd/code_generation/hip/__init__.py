"""HIP back end: replaces the OPS C printer for the explicit-solver path."""
from .target import HIPTarget, plan_problem, evaluate_parameters, initial_fields

__all__ = ["HIPTarget", "plan_problem", "evaluate_parameters", "initial_fields"]
