"""radiacuda_amd: MI355X (gfx950) implementation of Radia's interaction-matrix
assembly, relaxation solve and field evaluation, behind a Radia-compatible
Python API (``radiacuda_amd.rad``) and a C ABI (include/radhip.h)."""
__version__ = "0.1.0"
