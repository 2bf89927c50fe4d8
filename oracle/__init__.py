"""CPU oracle of the SBP-SAT Poisson path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, always as the checker; the product (sc_2022_amd) never does.
"""
