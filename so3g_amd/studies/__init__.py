"""Measurement studies built on the projection engine (precision, batching,
map-making)."""
