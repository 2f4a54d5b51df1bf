"""CPU restatement of the SIMPLE-TS time step and the reference build recipe
(test infrastructure only; see oracle.py)."""
