"""FP64 NumPy oracle of dLux's hot path -- test infrastructure only (see dlux_oracle)."""
