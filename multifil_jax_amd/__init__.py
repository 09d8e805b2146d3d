"""multifil_jax_amd -- MI355X-native per-timestep hot path of the multifil_jax
half-sarcomere simulator.

Public names mirror the reference package (multifil_jax/__init__.py:80-100)
for the parts this build provides: ``run``, ``SimulationResult``,
``SarcTopology``, ``get_skeletal_params`` / ``get_cardiac_params`` (plus the
invertebrate presets), ``StaticParams`` / ``DynamicParams``, ``Subpopulation``.  The step itself
runs in hand-written gfx950 HIP kernels (``csrc/``) behind a C ABI
(``include/multifil_amd.h``).
"""
from .params import (StaticParams, DynamicParams, Constants, get_skeletal_params,
                     get_cardiac_params, get_lethocerus_params, get_drosophila_params,
                     DYNAMIC_FIELDS)
from .topology import SarcTopology
from .subpopulation import Subpopulation


def __getattr__(name):
    # The simulation entry points load the device library on first use.
    if name in ("run", "SimulationResult", "METRIC_NAMES"):
        from . import simulation
        return getattr(simulation, name)
    raise AttributeError(name)


__all__ = ["run", "SimulationResult", "SarcTopology", "StaticParams", "DynamicParams",
           "Constants", "get_skeletal_params", "get_cardiac_params", "get_lethocerus_params",
           "get_drosophila_params", "DYNAMIC_FIELDS", "METRIC_NAMES", "Subpopulation"]
