"""FP64 CPU oracle for the split-step KDK Schroedinger-Poisson step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU restatement of the reference path).  The product
path in ``quantum-jax_amd`` never imports it.

The restatement follows the reference scripts line by line (file:line cited on
every function).  The reference itself (Python/JAX, ``jax[cuda12]==0.5.3``)
cannot run here: JAX is not installed and there is no network, so every
function is pinned by closed forms, invariants and brute force instead
(``tests/test_oracle.py``).  What only an import of the reference could pin --
its default-precision (FP32) rounding trajectory -- is explicitly unpinned.
"""

from .kdk import *  # noqa: F401,F403
