"""Checker for the Berry-quantity path (test infrastructure only; see wb_oracle)."""
