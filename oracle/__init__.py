"""CPU oracle for the LCODE 3D xi step (test infrastructure only)."""
