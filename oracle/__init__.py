"""CPU oracle for elemental generation (test infrastructure only; never imported by the product path)."""
