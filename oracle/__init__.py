"""The oracle: a CPU (numpy, complex128) restatement of the reference's Holstein TDVP
hot path.  TEST INFRASTRUCTURE ONLY -- only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import it; the product (kylin-v_amd/) never does.
See blocksparse.py, tdvp.py and model.py for the reference file:line each part follows.
"""
