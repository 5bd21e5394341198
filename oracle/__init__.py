"""Oracle package: the CPU checker (test infrastructure only)."""
