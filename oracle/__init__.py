"""CPU oracle for the FK forward-solve + likelihood path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  The product (gliomauq_amd) never imports it.
"""
