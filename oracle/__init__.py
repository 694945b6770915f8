"""Test-only CPU restatement of the reference's hot path (the oracle)."""
