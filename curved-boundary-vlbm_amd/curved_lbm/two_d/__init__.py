"""D2Q4x6 solvers."""
