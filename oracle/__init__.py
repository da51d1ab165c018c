"""CPU oracle for the pattern-matching path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only as the checker. See oracle.h for what each function
restates and tests/test_oracle.py for what pins it.
"""
