"""CPU oracle for the Newton-step hot path -- TEST INFRASTRUCTURE ONLY.

A plain-C FP64 restatement of the reference algorithm (``pfem_oracle.c``) plus a
small Python driver that reproduces ``newton!``/``solve!`` with direct sparse
solves.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package; the product package
``plasticity_amd`` never does.
"""
from .oracle import *  # noqa: F401,F403
