"""The volume ray march on the GPU: the reference's plume.render call surface over libplume_hip."""
