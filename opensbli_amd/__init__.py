"""opensbli_amd: OpenSBLI's RHS and Runge-Kutta time advance on MI355X (gfx950).

The problem-definition surface mirrors OpenSBLI's names (see
opensbli_amd.api); the compute path is libosbl_hip.so, reached through the
C-ABI declared in include/opensbli_amd.h.
"""
from . import native, problems  # noqa: F401

__all__ = ["native", "problems"]
