"""CPU oracle of the per-step loop.  Test infrastructure: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it, as the
checker and the CPU baseline.  The product never does."""
