import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "cuckoo-failure_amd"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")
    # a fresh checkout without a GPU: build the library and drivers first (hipcc cross-compiles);
    # where a GPU is present the prebuilt files are used as they are
    pkg = os.path.join(REPO, "cuckoo-failure_amd")
    built = all(os.path.exists(os.path.join(pkg, f)) for f in ("lib/libcuckoo_amd.so", "bin/cuckoo_fast",
                                                               "bin/cuckoo_fail", "bin/cf_post"))
    if not built and not os.path.exists("/dev/kfd"):
        subprocess.run(["make", "-s", "-j", str(min(16, os.cpu_count() or 1)), "-C", pkg], check=True)


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(REPO, "tests", "golden")
