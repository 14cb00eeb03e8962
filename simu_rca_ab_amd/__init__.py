"""MI355X-native RCA delay-and-sum I/Q beamformer (gfx950 HIP kernels behind a C-ABI)."""
__version__ = "0.1.0"
