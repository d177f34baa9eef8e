"""CPU oracle of the escape-time path: test infrastructure only (see oracle.py)."""
