"""capaz_amd: MI355X backend for capaz's explicit 3-D two-/three-phase
filtration step (and its energy equation), behind the C ABI of
include/capaz.h.  ``capaz_amd.capaz`` is the Python host binding;
``capaz_amd.workloads`` builds the seeded benchmark inputs;
``capaz_amd.deck`` reads the reference's input decks, runs its driver loop
and writes its Tecplot plots; ``capaz_amd.io`` reads and writes noise.dat and
state checkpoints."""

from . import capaz, deck, io, workloads  # noqa: F401

__all__ = ["capaz", "deck", "io", "workloads"]
