"""optlab for MI355X: the Monte Carlo path engine (`optlab.native`) and the host-side plan layer it reads
(`optlab.core`: markets, models, SVI/SSVI surfaces with Dupire local variance, the engine registry)."""
