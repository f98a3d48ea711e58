from mdcraft_amd.refapi import Boundary, Periodic
