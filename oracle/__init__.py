"""Oracle package: TEST INFRASTRUCTURE ONLY.

CPU FP64 restatement of mdff20's force/integration path (liboracle.so, see
mdff_oracle.c) and the runner of the reference build in _ref/ (refrun.py).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
as the checker; the product engine never imports it.
"""
