"""CPU oracle package (test infrastructure only; see hpcs_oracle.c)."""
