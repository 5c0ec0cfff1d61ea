"""CPU restatement of mujoco_warp's step — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker.  The product (mjx_warp_amd) never
imports it.
"""
