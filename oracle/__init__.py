"""FP64 CPU oracle for the multislice STEM path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy FP64, the reference's own CPU implementation of
the path: the MATLAB multicore engine

* ``GetPotential4AllSlice_multicore_lotabo_peng.m`` / ``..._corr.m`` (potential slices,
  transmission function), and
* ``cuda_STEM_CPU_Mcore_lobato_peng.m`` / ``..._corr.m`` (multislice propagation,
  annular detectors, focal-spread weighting, mid-slice outputs).

Only the test-suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline.
The product path (``totem_amd``) never imports, links or executes anything here.

Parity pins for each function are listed in the module docstrings and exercised
by ``tests/test_oracle.py``.
"""
