"""Shared test set-up: markers, paths and loaders for the oracle and the product."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
GROUPS = os.path.join(REPO, "dym_symanzik_action_amd", "groups")
for p in (REPO, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU or GPU case")


def group_file(name):
    return os.path.join(GROUPS, name + ".gz")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def ogroups(oracle):
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = oracle.OGroup(group_file(name))
        return cache[name]

    return get


@pytest.fixture(scope="session")
def dsa():
    import dym_symanzik_action_amd as m

    m.lib()
    return m


@pytest.fixture(scope="session")
def dgroups(dsa):
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = dsa.load_group(group_file(name))
        return cache[name]

    return get
