"""CPU oracle for the PatchMatch engine (test infrastructure only)."""
