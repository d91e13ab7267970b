"""CPU oracle of the Morphogenesis SPH step — test infrastructure only (see mg_oracle.cpp)."""
