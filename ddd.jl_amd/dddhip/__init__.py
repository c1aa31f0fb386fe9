"""MI355X-native force path for DDD.jl networks.

``dddhip.forces`` mirrors the reference's force-path API over the HIP C-ABI
library libdddhip.so (loaded on first use); ``dddhip.types`` mirrors its
parameter and network structs; ``dddhip.generate`` makes seeded synthetic
networks.
"""
from . import types, generate, forces  # noqa: F401

__all__ = ["types", "generate", "forces"]
