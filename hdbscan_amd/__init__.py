"""Batched, device-resident HDBSCAN for MI355X (gfx950) with a ROCm backend."""

from ._api import hdbscan

__all__ = ["hdbscan"]
