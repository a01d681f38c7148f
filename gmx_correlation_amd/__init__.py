"""MI355X-native all-pairs KSG mutual-information matrix (gmx_correlation hot path)."""
