from mdcraft_amd.refapi import NeibsList, VerletList
