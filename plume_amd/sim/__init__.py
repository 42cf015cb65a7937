"""The simulation substep on the GPU: the reference's plume.sim call surface over libplume_hip."""
