"""Standalone (HIP) backend: configuration and Python entry point."""

from .config import StandaloneBackendConfig
from .runtime import RingIsingAdjointBackend

__all__ = ["RingIsingAdjointBackend", "StandaloneBackendConfig"]
