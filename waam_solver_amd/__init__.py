"""waam_solver_amd — MI355X-native implementation of waam_twin's coupled physics step.

The step (LBM collide/stream, enthalpy thermal advection-diffusion with boundary losses,
enthalpy-porosity phase change, VOF advection with the wetting BC, CSF/Marangoni/buoyancy
forces, droplet deposition and pool telemetry) runs as hand-written HIP kernels for gfx950 in
libwaam_hip.so, called through a C ABI (include/waam_hip.h). ``WAAMTwin`` keeps the
reference's Python API.
"""

__version__ = "0.1.0"


def __getattr__(name):
    if name == "WAAMTwin":
        from .twin import WAAMTwin
        return WAAMTwin
    if name == "WAAMGrid":
        from .grid import WAAMGrid
        return WAAMGrid
    raise AttributeError(name)
