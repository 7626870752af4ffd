"""D3Q6x12 solvers."""
