"""FP64 CPU oracle of the tomography solve (test infrastructure only)."""
