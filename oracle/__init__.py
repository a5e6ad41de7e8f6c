"""The CPU oracle: a restatement of the reference's coupled substep.

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker and as the reported CPU baseline;
the product (eulersf_amd, libesf.so) never imports it.
"""
