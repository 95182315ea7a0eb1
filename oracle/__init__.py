"""The oracle: a CPU restatement of OpenSBLI's generated RHS/RK kernel sequence.

Test infrastructure only.  tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may load it, as the checker; the product never does.
"""
