"""CPU oracle (test infrastructure only): see uhm_oracle.c."""
