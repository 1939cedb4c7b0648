"""Backends of the ring-Ising workflow."""
