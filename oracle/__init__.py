"""CPU oracle of the soft-body FEM step (test infrastructure only)."""
