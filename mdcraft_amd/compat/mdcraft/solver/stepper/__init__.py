from mdcraft_amd.refapi import Stepper, Verlet
