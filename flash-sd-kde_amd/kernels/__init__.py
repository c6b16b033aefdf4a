"""Launch-level API of the hot path (module paths of the reference's kernels/ package)."""
