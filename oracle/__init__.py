"""CPU oracle for the gridding / degridding path -- TEST INFRASTRUCTURE ONLY.

This package restates crocodile's FP64 NumPy algorithms (``synthesis_ref``,
``wtowers_ref``) and a plain-C FP64 restatement of the per-visibility loops
(``cgrid.c`` reached through ``cgrid``).  It is the checker against which the
HIP implementation in ``crocodile_amd`` is compared, and the CPU baseline that
``bench.py`` times beside the GPU.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here.  The product package
``crocodile_amd`` never imports, links or executes this code.
"""
