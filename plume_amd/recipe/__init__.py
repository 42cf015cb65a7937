"""Recipes: the TOML scene description that drives a run (the reference's plume.recipe, restated for the
simulation and render keys the hot path reads)."""
