"""CPU oracle for the Lantern rendering hot path.

TEST INFRASTRUCTURE ONLY: the checker, never the product. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.
"""
