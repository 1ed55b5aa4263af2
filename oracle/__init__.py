"""CPU FP64 oracle of PHAFD's per-timestep hot path -- test infrastructure only.

Imported only by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py, as the checker.  The product (phafd_amd) never imports it.
"""
