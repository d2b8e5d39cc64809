"""ORACLE -- test infrastructure only (CPU restatements of the reference
algorithm).  Imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline/checker leg; never by the product package."""
