from mdcraft_amd.refapi import Threads
