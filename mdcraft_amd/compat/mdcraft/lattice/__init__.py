from mdcraft_amd.refapi import Domain
