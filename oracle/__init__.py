"""CPU oracle of the tracking path (test infrastructure only)."""
