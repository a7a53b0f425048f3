import os
import sys

# torch ships its own HIP runtime (libamdhip64, loaded by file name); when it
# is imported before librsc.so is loaded, the library binds to that same
# runtime (same soname) and both share one device context
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
