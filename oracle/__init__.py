"""CPU oracle for the GRG matmul -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
anything under oracle/, and only as the checker; the product never does.
"""
