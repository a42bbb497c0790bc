"""Test infrastructure: the float64 NumPy oracle of the SMI hot path (see smi_ref.py).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it."""
