"""CPU oracle (test infrastructure only): see evsl_oracle.h."""
