"""CPU oracle (test infrastructure only): see oracle/lbm_oracle.c."""
