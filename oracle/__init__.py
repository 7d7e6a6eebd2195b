"""CPU restatement of the reference foam step: test infrastructure only.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product never imports it.
"""
