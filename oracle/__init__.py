"""CPU oracle of the nlchains hot path -- test infrastructure only (see oracle.c)."""
