"""CPU oracle for the PEPS_MG multigrid path (test infrastructure only)."""
