"""FP64 host oracle (test infrastructure only)."""
