"""mjx_warp_amd: an MI355X-native batched physics step behind the mujoco_warp API.

Public surface of mujoco_warp/__init__.py:18-58: the same function names,
signatures and Model / Data field layout, computed by hand-written HIP
kernels for gfx950 (libmjx_warp_amd.so, called through a ctypes C ABI).
"""

from .api import collision as collision
from .api import com_pos as com_pos
from .api import com_vel as com_vel
from .api import crb as crb
from .api import euler as euler
from .api import factor_m as factor_m
from .api import forward as forward
from .api import fwd_acceleration as fwd_acceleration
from .api import fwd_actuation as fwd_actuation
from .api import fwd_position as fwd_position
from .api import fwd_velocity as fwd_velocity
from .api import implicit as implicit
from .api import kinematics as kinematics
from .api import make_constraint as make_constraint
from .api import mul_m as mul_m
from .api import nxn_broadphase as nxn_broadphase
from .api import passive as passive
from .api import primitive_narrowphase as primitive_narrowphase
from .api import rne as rne
from .api import rollout as rollout
from .api import sap_broadphase as sap_broadphase
from .api import solve as solve
from .api import solve_m as solve_m
from .api import step as step
from .api import step_fused as step_fused
from .api import StepGraph as StepGraph
from .api import transmission as transmission
from .api import xfrc_accumulate as xfrc_accumulate
from . import derivative as derivative
from .bridge import FunctionalStep as FunctionalStep
from .multi import MultiStep as MultiStep
from .bridge import Unroll as Unroll
from .io import get_data_into as get_data_into
from .tracing import EventTracer as EventTracer
from .tracing import event_scope as event_scope
from .io import is_sparse as is_sparse
from .io import make_data as make_data
from .io import put_data as put_data
from .io import put_model as put_model
from .test_util import benchmark as benchmark
from .types import BiasType as BiasType
from .types import ConeType as ConeType
from .types import Contact as Contact
from .types import Data as Data
from .types import DisableBit as DisableBit
from .types import DynType as DynType
from .types import GainType as GainType
from .types import GeomType as GeomType
from .types import IntegratorType as IntegratorType
from .types import JointType as JointType
from .types import Model as Model
from .types import Option as Option
from .types import SolverType as SolverType
from .types import Statistic as Statistic
from .types import TrnType as TrnType
