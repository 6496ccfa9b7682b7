"""CPU oracle (test infrastructure only; see ewo_oracle.h)."""
