from mdcraft_amd.refapi import Atoms
