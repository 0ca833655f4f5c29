"""CPU oracle of the MARL tick (test infrastructure; see marl_oracle.c)."""
