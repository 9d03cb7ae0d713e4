"""Drop-in for the reference's analytical_scattering_theories package."""
