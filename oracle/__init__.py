"""Test infrastructure: CPU float64 restatement of the reference time loop (the parity oracle)."""
