"""Test infrastructure: the FP64 CPU oracle (see gbg_oracle.py).  Never part
of the product path."""
