"""srf_amd — MI355X-native projector hot path of the SRF PET reconstruction
framework: Siddon and TOF tube-of-response projection / backprojection, the
efficiency-map backprojection and the list-mode MLEM/OSEM update, as hand-written
HIP kernels for gfx950 behind a C ABI (include/srfx.h, libsrfx.so)."""

__version__ = "0.1.0"
