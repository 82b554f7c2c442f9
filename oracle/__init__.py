"""Oracle for the k-space first-order PSTD time loop -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may import
this package, and only as the parity reference or the CPU baseline.  The product
(k-wave-julia_amd/) never imports it.

The reference (KWave.jl) is Julia; no Julia toolchain exists in this image,
so there is no oracle/_ref build: the oracle is a from-source FP64 NumPy/SciPy
restatement (kwave_oracle.py) plus the exact per-mode recurrence
(closed_form.py) that pins it.
"""
