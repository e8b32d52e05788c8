"""CPU oracle of the ColBERT late-interaction search path (test infrastructure only)."""
