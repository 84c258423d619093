"""Rule decorators (framework/src/oncolytica/cpu/_decorators.py:23-59): they
only tag methods; Engine.load_model discovers rules by the tag."""
from __future__ import annotations

from typing import Callable, Optional


def tissue_rule(_func: Optional[Callable] = None) -> Callable:
    def deco(f):
        f._rule_type = "tissue"
        return f
    return deco(_func) if _func is not None else deco


def chemistry_rule(iterations: int = 1) -> Callable:
    def deco(f):
        f._rule_type = "chemistry"
        f._iterations = iterations
        return f
    return deco


def cell_rule(_func: Optional[Callable] = None) -> Callable:
    def deco(f):
        f._rule_type = "cell"
        return f
    return deco(_func) if _func is not None else deco


def metric_rule(_func: Optional[Callable] = None) -> Callable:
    def deco(f):
        f._rule_type = "metric"
        return f
    return deco(_func) if _func is not None else deco
