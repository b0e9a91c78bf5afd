"""Fourier-neural-operator warm start for steady-state sweeps (an addition of this package)."""
