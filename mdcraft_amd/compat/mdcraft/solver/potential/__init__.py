from mdcraft_amd.refapi import (EAM, LJs, MLIP4, DeepModelingPotential, LJsCut, MaiseBPNNPotential, Pair, Potential,
                                PotentialTernary)
