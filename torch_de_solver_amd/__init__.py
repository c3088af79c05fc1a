"""MI355X-native residual/loss engine for TEDEouS-style solvers ('NN' and 'mat' modes).

Public API mirrors the reference's (tedeous.data, tedeous.model,
tedeous.optimizers.optimizer, tedeous.models, tedeous.device).
"""

from .data import Conditions, Domain, Equation
from .device import check_device, device_type, solver_device
from .model import Model
from .models import mat_model
from .optimizers import Optimizer
from .solution import Operator_bcond_preproc, Solution

__all__ = ['Conditions', 'Domain', 'Equation', 'Model', 'Optimizer', 'Solution',
           'Operator_bcond_preproc', 'mat_model', 'solver_device', 'check_device',
           'device_type']
