"""MI355X-native size-bounded community search (SCS)."""
