"""piqmc -- MI355X (gfx950) implementation of hadsed/pathintegral-qmc's sweeps.

Drop-in modules with the reference's names and signatures:
    piqmc.qmc.QuantumAnneal, piqmc.qmc.QuantumAnneal_parallel
    piqmc.sa.Anneal, piqmc.sa.ClassicalIsingEnergy
    piqmc.tools.GenerateNeighbors, piqmc.tools.bits2spins, piqmc.tools.spins2bits
Batched device API: piqmc.device.Problem.

Everything computes through libpiqmc_amd.so (hand-written HIP kernels); the
import fails if that library is not built.
"""
from . import _lib  # noqa: F401  (loads libpiqmc_amd.so or raises ImportError)
from . import qmc, sa, tools  # noqa: F401

__all__ = ["qmc", "sa", "tools"]
