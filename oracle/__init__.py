"""The oracle: FP64 CPU restatement of the reference algorithm (test infrastructure only)."""
