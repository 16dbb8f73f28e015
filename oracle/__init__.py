"""CPU restatement of the reference pixel pipeline — test infrastructure only."""
