"""FP64 CPU oracle of se2wave's time step (test infrastructure only)."""
