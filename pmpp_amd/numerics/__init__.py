from .fft import rfftn, irfftn, spectral_force_multiply  # noqa: F401
