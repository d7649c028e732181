"""wavesim_amd: wavecore's time step (FV + SGN) on MI355X, behind wavecore's Grid/State/Solver API.

The compute runs in hand-written HIP kernels for gfx950 in ``libwavecore_hip.so`` (C ABI in
``include/wavecore_hip.h``); this package is the host-side mirror of the reference's API.
"""

from .config import (  # noqa: F401
    Breaking,
    EddyViscosity,
    Dispersion,
    DispersionStats,
    Limiting,
    LinearWave,
    Order,
    Sponge,
    WaveMaker,
    linear_wave,
    manning_factor,
)
from .grid import GHOST, G, H_DRY, THREADED_CELLS, Bathymetry, Grid, State  # noqa: F401
from .solver import (  # noqa: F401,E402
    PRECISION_FP32,
    Batch,
    PRECISION_FP64,
    DeviceError,
    Relaxation,
    Sampler,
    Solver,
    Stats,
    StepWatch,
    stage_step,
)
