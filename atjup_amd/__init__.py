"""MI355X-native implementation of the ATJUP time step (RogerGrundmann/ATJUP)."""
