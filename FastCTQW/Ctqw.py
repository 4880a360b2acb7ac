"""Alias of fastctqw_amd.Ctqw (reference: FastCTQW/Ctqw.py)."""
from fastctqw_amd.Ctqw import Ctqw  # noqa: F401
