"""MI355X-native SBP-SAT hot path (placeholder during bring-up)."""
