"""CPU FP64 oracle for parity tests (test infrastructure; never imported by the library)."""
