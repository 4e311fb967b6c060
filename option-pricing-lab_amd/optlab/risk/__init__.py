"""Risk layer restated for the MI355X engine: the risk grid (packages/risk/src/optlab/risk/grid.py) and implied-vol
surfaces held on the device, whose bumped local-variance tables the engine builds itself
(packages/risk/src/optlab/risk/surfaces.py)."""
