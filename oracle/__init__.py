"""Oracle for the fftset transform / fast-convolution path.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` as the checker. The product library
(``fftset_amd``) never imports it.

Two oracles live here:

* ``Oracle`` -- ctypes binding of ``fftset_oracle.c``, a CPU FP32 restatement
  of the reference's scalar path (planner, twiddles, Stockham / DIF / DIT
  passes, FREQOFFSETREAL and COMPLEX modulators). It stands in for the
  reference C library, which cannot be built in this image (its SIMD/allocator
  dependency ``cop`` is absent).
* ``fp64`` -- closed-form FP64 references (DFT, odd-frequency DFT, cyclic and
  negacyclic convolution, adjoint gradients).
"""
from .fp64 import *  # noqa: F401,F403
from .oracle import Oracle, OraclePlan, load_oracle, oracle_available, ORACLE_SO  # noqa: F401
