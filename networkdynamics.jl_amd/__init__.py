"""MI355X-native network right-hand side of NetworkDynamics.jl.

    import networkdynamics_jl_amd as ND
    g  = ND.barabasi_albert(1_000_000, 2, seed=1)
    nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge(),
                    execution=ND.HIPExecution(True), aggregator=ND.SegmentedAggregator)
    nw(du, u, p, t)          # du, u, p: float64 torch tensors on the HIP device

The directory is named `networkdynamics.jl_amd`; it is importable under the
module name `networkdynamics_jl_amd` through `load()` in __graft_entry__.py.
"""
from . import _abi
from . import library as Lib
from . import workloads
from .ensemble import Ensemble, tsit5_per_instance
from .fixpoint import FixpointSolver, swing_load_grid, isfixpoint_residual
from . import fixpoint
from ._abi import ArgumentError, DeviceError, NotConvergedError, UnsupportedError
from .components import (AntiSymmetric, DeviceFunction, Diagonal, Directed, EdgeModel, EIndex, FeedForward,
                         Fiducial, I, NoFeedForward, PureFeedForward, PureStateMap, StateMask, Symmetric,
                         UniformScaling, VertexModel, VIndex, fftype, hasff)
from .graphs import (SimpleDiGraph, SimpleGraph, barabasi_albert, complete_graph, cycle_graph, erdos_renyi,
                     path_graph, random_geometric, star_digraph, watts_strogatz)
from .network import (Aggregator, ExecutionStyle, HIPExecution, Network, SegmentedAggregator, usebuffer)


def dim(nw):
    return nw.dim


def pdim(nw):
    return nw.pdim


__all__ = [n for n in dir() if not n.startswith("_")]
from . import warmstart
from . import experiments
from . import ensemble_tune
from . import linear_analysis
from . import sparsity
from . import dae
from .initialization import NetworkInitError, find_fixpoint
