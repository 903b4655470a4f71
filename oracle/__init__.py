"""CPU FP64 oracle (test infrastructure only; see gato_oracle's header)."""
