from mdcraft_amd.refapi import (EAM, EAMAlloy, LJs, MLIP4, DeepModelingPotential, LJsCut, MaiseBPNNPotential, Potential,
                                PotentialTernary)
from mdcraft_amd.refapi import TabulatedPair as Pair
