"""Operator modules of the coupled step (reference physics/*)."""
