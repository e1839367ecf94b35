"""CPU oracle of the Disco timestep (test infrastructure; see disco_oracle.h)."""
