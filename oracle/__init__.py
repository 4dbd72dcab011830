"""CPU checker of the Vext path (test infrastructure only; see vext_oracle.py)."""
