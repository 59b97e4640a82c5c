"""CPU oracle for the NeuralODE forward + InterpolatingAdjoint hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline).  The product path (the HIP library
behind ``include/node_hip.h`` and the ``diffeqflux_hip`` host package) never
imports, links or calls anything here; it fails loudly when the HIP library is
missing.

The reference's arithmetic for this path lives in third-party Julia packages
that are absent from the reference tree (no Manifest; versions pinned only by
``Project.toml`` [compat] lower bounds):

* OrdinaryDiffEq 7 (``Tsit5``: tableau, embedded error, PI controller, initial
  dt, tstops, dense output, saveat) -- called at ``src/neural_de.jl:55-56``;
* SciMLSensitivity >= 7.112 (``InterpolatingAdjoint(autojacvec=ZygoteVJP())``)
  -- selected at ``src/neural_de.jl:57``;
* Lux >= 1.29 / LuxLib >= 1.12.1 (``Chain``/``Dense`` forward and pullback).

This package restates their published algorithms in numpy (FP64 by default;
FP32 when timed as the CPU baseline).  Every function names what pins it in
the module docstrings; ``tests/test_oracle_*.py`` hold the pins.

Modules:
  tableau   -- Tsit5 coefficients (Tsitouras 2011) and dense-output weights
  mlp       -- Lux Chain of Dense layers: forward, VJP, flat parameter layout
  tsit5     -- the OrdinaryDiffEq integrator loop (adaptive + fixed dt)
  adjoint   -- SciMLSensitivity InterpolatingAdjoint (continuous adjoint)
  ffjord    -- FFJORD augmented RHS (Hutchinson / exact trace) and log-likelihood
  problems  -- the five workload definitions (synthetic data generators)
"""
