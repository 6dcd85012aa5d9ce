"""ORACLE — test infrastructure only (CPU restatement of the reference's
device-load path).  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import it, and only as the checker."""
