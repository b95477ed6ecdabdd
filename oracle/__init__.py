"""CPU oracle for the Monte Carlo fault-tree path (test infrastructure only).

Only the tests, the smoke check and the benchmark's CPU-baseline leg may
import this package, and only as the checker.
"""
