"""Import path of the float64 NumPy flow oracle (oracle/flow_oracle.py) for the tests."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import flow_oracle  # noqa: E402,F401
