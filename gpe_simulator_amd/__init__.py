"""MI355X-native split-step Gross-Pitaevskii engine (libgpe_hip.so + host API).

Drop-in for the propagation path of PoyntingZhang/GPE_simulator: the same
schema-2 YAML, `stage`, `make_real_propagator` / `make_imag_propagator` and
`launch_stage`, with all device work in hand-written HIP for gfx950.
"""

__version__ = "0.1.0"
