import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libpvl_hip.so")
    # The oracle (test infrastructure) is plain C; build it if this tree has not yet.
    orc = ROOT / "oracle" / "liborc.so"
    if not orc.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liborc.so"], check=True)
