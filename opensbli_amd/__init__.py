"""MI355X-native OpenSBLI explicit-solver hot path."""
