"""The multi-name inputs of the path engine's plans: baskets with a correlation root and the local-correlation
mixer (packages/structured/src/optlab/structured/multiasset.py, localcorr.py)."""
