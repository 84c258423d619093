"""CompilationError: raised when a rule cannot be translated (same shape as
framework/src/oncolytica/core/utils/_errors.py:9-19)."""
from __future__ import annotations

import ast


class CompilationError(Exception):
    def __init__(self, message: str, node: ast.AST | None = None) -> None:
        self.raw_message = message
        self.node = node
        self.lineno = getattr(node, "lineno", None) if node is not None else None
        super().__init__(f"Line {self.lineno or '?'}: {message}")
