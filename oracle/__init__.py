"""CPU oracle for the half-sarcomere hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import or call anything here, and only as the
checker or the timed CPU baseline; the product package never does.

Contents
--------
``mf_oracle``   FP64 NumPy restatement of the reference's per-timestep path
                (multifil_jax/timestep.py, kernels/*.py, simulation.py:633-747,
                metrics_fn.py), one function per reference function, each
                citing the reference file:line it follows.
``philox``      Philox4x32-10 (SC'11), regenerates the device's uniforms.
``cpu_port.c``  OpenMP C restatement (FP32 bulk, same caps) used as the timed
                CPU baseline; built by ``oracle/Makefile``.

Pinning (what checks each function; see tests/test_oracle_*.py):
expm -> scipy.linalg.expm, semigroup, 2-state closed form; rates -> detailed
balance identities and the docstring values of the reference; forces ->
torch float64 autograd of an independently written energy; JVP -> central
differences; preconditioner -> closed-form Green's functions; Newton-CG ->
an exact sparse Newton solve; sampling / nearest site -> brute force.
The reference has no tests or golden vectors of its own, and JAX is not
installed here, so no function is pinned against reference outputs.
"""
