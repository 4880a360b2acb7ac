"""Alias of fastctqw_amd.State (reference: FastCTQW/State.py)."""
from fastctqw_amd.State import State  # noqa: F401
