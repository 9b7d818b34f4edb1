"""MI355X-native explicit time step for EQdyna-style dynamic rupture."""
