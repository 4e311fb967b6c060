"""Host-side semantics the path engine's plans are built from: the engine registry (engines.py), markets,
models and time grids, SVI / SSVI surfaces and their Dupire local variance, closed forms."""
