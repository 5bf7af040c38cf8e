"""Back ends that turn an OpenSBLI algorithm into a running solver."""
