"""MI355X-native FastCTQW: CTQW simulation with a gfx950 matrix exponential.

Public surface mirrors the reference package (FastCTQW/__init__.py:3):
``Ctqw``, ``Operator``, ``State`` modules and the ``fastexpm`` subpackage.
Importing this package does not load the GPU library; ``fastexpm`` does.
"""
from . import Ctqw, Operator, State  # noqa: F401

__version__ = "2026.1.0"
