"""The backend plugin protocol, unchanged from
framework/src/oncolytica/core/runtime/_backend.py:6-36; HipBackend
(oncolytica_amd/hip/_backend.py) implements it."""
from __future__ import annotations

from typing import Any, Protocol, runtime_checkable


@runtime_checkable
class ISimulationBackend(Protocol):
    def compile(self, sim_instance: Any) -> None:
        """Compile / initialise for one simulation instance (once, after load_model)."""

    def run_step(self, collect_metrics: bool = False) -> None:
        """Execute one simulation step."""

    def sync_to_host(self) -> None:
        """Device -> host state."""

    def sync_to_device(self) -> None:
        """Host -> device state."""

    def get_metrics(self) -> Any:
        """Last updated metrics."""
