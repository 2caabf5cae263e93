"""TEST INFRASTRUCTURE ONLY.

The CPU oracle for the domain-wall hot path: a C++ restatement of the
reference operator and solver (dwf_oracle.cpp, reached through oracle.py),
an analytic free-field propagator (freefield.py) and an independent numpy
statement of the 5D operator (dense.py).  Only tests/, the smoke entry and
the CPU-baseline leg of the benchmark may import this package; the product
library never does.
"""
