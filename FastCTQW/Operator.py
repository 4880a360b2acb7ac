"""Alias of fastctqw_amd.Operator (reference: FastCTQW/Operator.py)."""
from fastctqw_amd.Operator import Operator  # noqa: F401
