"""CPU oracle for the warp-preconditioners hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
import this package, and only as the checker.  The product (the HIP library and
the ``warp_preconditioners`` package) never imports it.

``oracle.cpu`` wraps ``_build/liboracle.so`` (built from ``wpc_oracle.cpp`` by
``oracle/Makefile``) with numpy-level pipelines that mirror the reference's host
logic (fsai.py:222-344, fsai_reuse.py:157-245, block_ilu.py:225-450,
squared_laplacian.py:103-218).
"""
