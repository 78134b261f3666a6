"""optifax_amd: MI355X (gfx950) engine for OptiFax's chi(2) split-step crystal
solver and synchronously pumped OPO round-trip loop.

The compute path is libofx_hip.so (hand-written HIP, built in-tree); this
package is the host interface: input producers (physics), the solver handle
with the reference's signature (solver.OPOmexBatch) and the device-resident
round-trip engine (solver.OPOBatch).
"""
from . import physics  # noqa: F401
from ._lib import LIB_PATH, OfxError  # noqa: F401

__all__ = ["physics", "solver", "LIB_PATH", "OfxError"]


def __getattr__(name):
    if name == "solver":
        from . import solver
        return solver
    raise AttributeError(name)
