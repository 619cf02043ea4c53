"""CPU oracle of ConSGA (test infrastructure only; see oracle/oracle.h)."""
