"""CPU oracle of the path (test infrastructure only): oracle.c restates the
reference's algorithms; oracle.py wraps it."""
