"""fastexpm: GPU matrix exponential classes (same names as the reference's
FastCTQW/fastexpm/__init__.py:1)."""
from ._wrapper import ExpMatBatch, ExpMatFloat32, ExpMatFloat64, ExpMatComplex64, ExpMatComplex128

__all__ = ["ExpMatFloat32", "ExpMatFloat64", "ExpMatComplex64", "ExpMatComplex128", "ExpMatBatch"]
