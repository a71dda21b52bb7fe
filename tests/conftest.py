import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU test")


SITE = (15.65, 46.56, 1)          # lon, lat, tz of the benchmark configurations
CLI_SITE = (15.592, 46.543, 1)    # reference CLI defaults (main.cpp:74-81)
