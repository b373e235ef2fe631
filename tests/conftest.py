import os
import sys

# The oracle's OpenMP loops: sleep instead of spinning while waiting, so that
# a run sharing the host's cores (other test processes, a fixture run) does
# not stall; read by libgomp when the oracle library loads.
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libspectral_hip.so")
    config.addinivalue_line("markers", "slow: CPU test that runs for tens of seconds")
