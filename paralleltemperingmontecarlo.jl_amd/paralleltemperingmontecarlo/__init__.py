"""MI355X-native parallel-tempering Monte Carlo engine with the
ParallelTemperingMonteCarlo.jl API for the PT hot path.

The Monte-Carlo cycles (displacement moves, incremental energies,
Metropolis, replica exchange, adaptation, sampling) run in HIP kernels in
lib/libptmc.so; this package is the host-side mirror of the reference's
Julia types and driver.
"""
from .api import (  # noqa: F401
    AngularType3,
    CubicBC,
    Config,
    ELJPotential,
    ELJPotentialB,
    ELJPotentialEven,
    EmbeddedAtomPotential,
    EngineOptions,
    MCParams,
    MCState,
    NeuralNetworkPotential,
    RuNNerPotential2Atom,
    RadialType2a,
    AngularType3a,
    nnp2_energies,
    NPT,
    NVT,
    NNVT,
    LookupTablePotential,
    Output,
    PTEngine,
    PotentialPlugin,
    RadialType2,
    RectangularBC,
    RhombicBC,
    RuNNerPotential,
    SphericalBC,
    TempGrid,
    check_e_bounds,
    forward_pass,
    get_volume,
    kB,
    ptmc_run,
    ptmc_run_batch,
    read_lookuptable,
    read_runner_scaling,
    read_runner_weights,
)
from ._lib import PTMCError, load as load_library  # noqa: F401
from .multihist import (  # noqa: F401
    multihistogram, multihistogram_NPT, multihistogram_npt_batch, multihistogram_NVT, multihistogram_nvt_batch,
    postprocess)

AtoBohr = 1.8897259886
