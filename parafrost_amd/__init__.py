"""MI355X-native GPU simplifier (SIGmA inprocessing engine) for SAT solvers.

The product is the C-ABI library ``lib/libpfsimp.so`` (include/pfsimp.h),
built from the HIP sources in ``csrc/``; ``simplifier`` is its Python
binding, ``gen`` the seeded workload generators and ``pfd`` the dump format.
"""
from . import pfd  # noqa: F401

__all__ = ["pfd", "gen", "simplifier"]
