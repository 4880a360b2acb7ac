"""Alias of fastctqw_amd.fastexpm (reference: FastCTQW/fastexpm/__init__.py:1)."""
from fastctqw_amd.fastexpm import (  # noqa: F401
    ExpMatComplex64,
    ExpMatComplex128,
    ExpMatFloat32,
    ExpMatFloat64,
)
