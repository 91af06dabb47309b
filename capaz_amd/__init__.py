"""capaz_amd: MI355X backend for capaz's explicit 3-D two-/three-phase
filtration step (and its energy equation), behind the C ABI of
include/capaz.h.  ``capaz_amd.capaz`` is the Python host binding;
``capaz_amd.workloads`` builds the seeded benchmark inputs."""

from . import capaz, workloads  # noqa: F401

__all__ = ["capaz", "workloads"]
