import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Under pytest-xdist every worker loads the oracle with its own OpenMP team.
# Teams of one thread per core in every worker oversubscribe the host, and the
# oracle's short parallel loops then wait on descheduled threads (a 1 s test
# took minutes): give each worker its share of the cores unless the
# environment already fixes the count.  libgomp reads the variable when the
# oracle library is first loaded, which is after this file runs.
_workers = os.environ.get("PYTEST_XDIST_WORKER_COUNT")
if _workers and "OMP_NUM_THREADS" not in os.environ:
    _cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(max(1, _cores // max(1, int(_workers))))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
