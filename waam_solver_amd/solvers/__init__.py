"""Step orchestration."""
