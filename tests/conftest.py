import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.load()
    return binding


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_kernels.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def config_driver(tmp_path_factory):
    """tests/cpp/config_driver.cpp (the include/icarion_config.hpp helpers over
    the reference's configuration files) built with g++; a function
    run(config, collision=None) -> the printed descriptors as a dict."""
    import json
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("cfg") / "config_driver")
    subprocess.run([gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "config_driver.cpp"),
                    "-o", exe], check=True)
    cfg = os.path.join(ROOT, "tests", "golden", "config")

    def run(config, collision=None):
        args = [exe, os.path.join(cfg, config), os.path.join(cfg, "species_database_v1.json"), cfg]
        if collision:
            args.append(collision)
        return json.loads(subprocess.run(args, check=True, capture_output=True, text=True).stdout)

    run.exe = exe
    return run
