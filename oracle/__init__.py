"""CPU restatement of the reference solvers: the checker of the HIP solver.

Test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's CPU
baseline).  See binding.py.
"""
