"""FP64 CPU restatement of the reference (test infrastructure only)."""
