"""CPU oracle of the compact FD core (test infrastructure only)."""
