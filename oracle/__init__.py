"""CPU oracle for the HIP phaseless-AFQMC path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product package never does.
"""
