"""The oracle: a CPU restatement of NumFast's relational hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this package, and only as the
checker (or the timed CPU baseline) -- never as a compute path of
numfast_amd.

Two layers:

* ``nf_oracle.c`` (built to ``libnf_oracle.so``) restates the reference
  crate numfast-native function by function with the same C-ABI symbol
  names, so one ctypes harness drives both it and libnumfast_hip.so;
* ``executor.py`` restates, in NumPy, the executor-level semantics that sit
  above the ABI: NULL compaction, group-by with sum/count/min/max/mean
  (ColumnCarry), unique with validity, NULL-last multi-key sort, the
  filter/derive/sort/limit pipeline, and an exact-rational float64 sum.

How each function is pinned (golden vectors, NumPy library routines, brute
force) is stated in its docstring and in tests/test_oracle_*.py.
"""
