import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)



@pytest.fixture(scope="session")
def ctx():
    from mdcraft_amd import engine
    c = engine.Context(0)
    yield c
