"""MI355X-native solver core for SimInf-style stochastic compartment
models: the ssm/mssm/aem solvers, scheduled events and trajectory
writeback as hand-written HIP for gfx950, behind the reference's
SimInf_solver_args C boundary (include/siminf_hip.h)."""
from .model import SimInfModel, SimInfError, make_events, run, trajectory
from .models import SIR, SEIR, SIS, SISe, SISe3, SISe_sp, SISe3_sp, SIR_mparse
from .spatial import distance_matrix, ldata_sp, add_spatial_coupling_to_ldata
from .modelparser import mparse
from .particle_filter import pfilter, PFilterResult
from .plan import Plan, Batch
from .punchcard import punchcard
from .abc_smc import abc, ABCResult
from .pmcmc import pmcmc, PMCMCResult
from .summaries import SummaryError
from . import _native

__all__ = ["SimInfModel", "SimInfError", "make_events", "run", "trajectory",
           "SIR", "SEIR", "SIS", "SISe", "SISe3", "SISe_sp", "SISe3_sp", "SIR_mparse", "mparse",
           "distance_matrix", "ldata_sp", "add_spatial_coupling_to_ldata",
           "pfilter", "PFilterResult", "Plan", "Batch", "SummaryError", "punchcard", "abc", "ABCResult",
           "pmcmc", "PMCMCResult"]
