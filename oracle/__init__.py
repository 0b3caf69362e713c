"""Test infrastructure: the checker for the HIP hot path.

``oracle/`` holds a CPU restatement in plain C of the reference FIM dynamics
hot path (``fim_oracle.c``), the recipe that builds the reference Fortran
itself into ``oracle/_ref`` (``build_ref.sh``), and Python drivers for both.
Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker or the
timed CPU baseline - never as the product path.
"""
