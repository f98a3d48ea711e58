from mdcraft_amd.refapi import ISolver, Multi, Single, Ternary
from mdcraft_amd.refapi import _EngineSolver as Pair
from .boundary import Boundary
from .thermostat import Thermostat
