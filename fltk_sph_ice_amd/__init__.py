"""MI355X-native particle hot path of the SPH/ice simulator: neighbour grid, WCSPH, PBF, shape matching and the
surface extraction (anisotropic kernels, marching cubes) as hand-written HIP kernels for gfx950 behind the
reference's extern "C" particle ABI.

The native library is fltk_sph_ice_amd/librxcu.so (built in-tree by fltk_sph_ice_amd/build.py); every entry
point here calls into it and raises when it is missing.
"""
from ._abi import check, load, rxParticleCell, rxSimParams, rxSphEnv  # noqa: F401

__all__ = ["load", "check", "rxSimParams", "rxParticleCell", "rxSphEnv"]
