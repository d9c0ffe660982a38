"""CPU checkers for tests and benchmarks (test infrastructure only)."""
