from mdcraft_amd.refapi import Compound, Langevin, Thermostat
