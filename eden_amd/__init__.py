"""eden_amd -- MI355X (gfx950) per-timestep simulation loop for EDEN NeuroML models.

The HIP library ``libedenhip.so`` (built in-tree by ``python -m eden_amd.build``) does the
simulation; this package builds instantiated models, drives the library through its C-ABI
and mirrors EDEN's Python entry point ``runEden``.
"""
from .model import Model, Simulation, EdenHipError  # noqa: F401
