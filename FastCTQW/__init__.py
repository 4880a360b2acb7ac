"""Drop-in import path of the reference package (``FastCTQW``).

Re-exports the MI355X implementation in ``fastctqw_amd`` under the module
names of the reference (FastCTQW/__init__.py:3), so that
``from FastCTQW.Ctqw import Ctqw`` and ``from FastCTQW.fastexpm import
ExpMatComplex64`` work unchanged.
"""
from fastctqw_amd import Ctqw, Operator, State  # noqa: F401
from fastctqw_amd import __version__  # noqa: F401
