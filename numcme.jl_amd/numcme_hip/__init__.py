"""numcme_hip — MI355X-native finite-state-projection engine with NumCME.jl's API.

Host-side mirror of NumCME.jl's FSP interface over the C ABI of ``libnumcme_hip.so``
(``include/numcme_hip.h``). The state space, generator, integrators and adapters run on
the GPU; there is no CPU fallback.
"""
from ._lib import ArgumentError, NumCmeHipError, device_name, expm_dense
from .expr import ifelse
from .fspmatrix import FspMatrixSparse, matvec, matvec_, matvecadd_
from .fspvector import FspVectorSparse, nnz
from .model import (CmeModel, CmeModelWithSensitivity, JointTimeVaryingPropensity, Propensity,
                    SeparableTimeVaryingPropensity, StandardTimeInvariantPropensity, istimeseparable,
                    istimevarying, propensity)
from .solve import RK45, FspOutputSliceSparse, FspOutputSparse, KrylovExpv, solve
from .adaptive import (AdaptiveFspSparse, LearnedAdapter, RStepAdapter, SelectiveRStepAdapter, adapt_, init_,
                       solve_adaptive_batch)
from .batch import BatchFsp
from .mlp import AdapterMlp
from .learned import TrainingRecords, boundary_features, train_adapter
from .sens import (AdaptiveForwardSensFspSparse, ForwardSensFspInitialConditionSparse, ForwardSensFspMatrixSparse,
                   ForwardSensFspOutputSliceSparse, ForwardSensFspOutputSparse, ForwardSensRStepAdapter,
                   forwardsens_initial_condition)
from .statespace import (StateSpaceSparse, deleteat, expand, get_sink_connectivity, get_sink_count,
                         get_state_connectivity, get_state_count, get_statedict, get_states, get_stoich_matrix)
from . import expr as nmath

__all__ = [
    "ArgumentError", "NumCmeHipError", "device_name", "expm_dense", "ifelse", "nmath",
    "FspMatrixSparse", "matvec", "matvec_", "matvecadd_", "FspVectorSparse", "nnz",
    "CmeModel", "CmeModelWithSensitivity", "JointTimeVaryingPropensity", "Propensity",
    "SeparableTimeVaryingPropensity", "StandardTimeInvariantPropensity", "istimeseparable", "istimevarying",
    "propensity", "RK45", "FspOutputSliceSparse", "FspOutputSparse", "KrylovExpv", "solve",
    "AdaptiveFspSparse", "RStepAdapter", "SelectiveRStepAdapter", "adapt_", "init_",
    "BatchFsp", "AdaptiveForwardSensFspSparse", "ForwardSensFspInitialConditionSparse", "ForwardSensFspMatrixSparse",
    "ForwardSensFspOutputSliceSparse", "ForwardSensFspOutputSparse", "ForwardSensRStepAdapter",
    "forwardsens_initial_condition",
    "StateSpaceSparse", "deleteat", "expand", "get_sink_connectivity", "get_sink_count",
    "get_state_connectivity", "get_state_count", "get_statedict", "get_states", "get_stoich_matrix",
]
