"""MI355X-native Rayleigh-Chebyshev lowest-eigenpair solver."""
