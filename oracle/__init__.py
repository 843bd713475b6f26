"""CPU oracle for DUST's hot path -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the reference results are compared against; the product (dust_amd,
libdust_gpu.so) never imports it.
"""
