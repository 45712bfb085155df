"""mcd — MI355X-native electron-trajectory transport for MC-discrete.

The package mirrors the reference's host interface for the scattering loop
(MC/tools/multiScatter.py, MC/tools/fileTools.py, MC/doScatter.py) and runs
the loop itself in hand-written gfx950 HIP kernels through libmcd.so.
"""
__version__ = "0.1.0"
