"""CPU FP64 oracle of the cloth substep loop: test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
as the checker; the product library never loads it.
"""
