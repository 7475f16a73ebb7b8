"""kitbase.jl_amd -- MI355X (gfx950) discrete-velocity time stepper for KitBase.jl's hot loop.

The directory name is not a Python identifier; load it as the package `kitbase_amd`:

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "kitbase_amd", "kitbase.jl_amd/__init__.py", submodule_search_locations=["kitbase.jl_amd"])
    kb = importlib.util.module_from_spec(spec); sys.modules["kitbase_amd"] = kb; spec.loader.exec_module(kb)

Contents: solver.py (SolverSet, init_fvm, timestep, DeviceSolve, reconstruct/evolve/update/solve),
phase.py, geometry.py, theory.py (setup restatements), _lib.py (ctypes over libkbgpu.so),
serving.py (request loop: admission by footprint, continuous batching, p50/p99), autotune.py
(schedule tuner), configs.py
(the workloads), csrc/ (HIP kernels and the C ABI), build.py (hipcc recipe).
"""
from . import _lib, autotune, configs, fsm, geometry, mesh, phase, serving, theory  # noqa: F401
from .solver import (Context, ControlVolumes, DeviceSolve, Interfaces, SolverSet, batch_step, batch_step_cfl,  # noqa: F401
                     batch_timestep, default_context, evolve, init_fvm, make_desc, read_dict, reconstruct, solve,
                     solve_bytes, timestep, update)

__version__ = "1.0.0"
