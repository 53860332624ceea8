"""pytest configuration: the `gpu` marker and import aliases.

The product package lives in the directory `pentadiagonal-solver_amd/`, whose
name is not a Python identifier; it is imported here under the alias
`pentadsolver_amd`.  The CPU oracle (test infrastructure) is imported as
`oracle`.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pentadiagonal-solver_amd")


def _alias(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(path, "__init__.py"), submodule_search_locations=[path])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


_alias("pentadsolver_amd", PKG)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = os.path.join(GOLDEN, "reference_fixtures")
FIXTURE_NAMES = ["one_dim_small_solve0", "one_dim_large_solve0",
                 "two_dim_small_solve0", "two_dim_small_solve1",
                 "two_dim_large_solve0", "two_dim_large_solve1"]


# The reference's three 3-D fixtures (three_dim_large_solve{0,1,2}, 32^3) are
# missing from its checkout; tests/golden/gen_fixtures.py restates its
# generator and writes them (tests/test_fixture_generator.py pins the
# generator against the six shipped fixtures).
GENERATED_3D = ["three_dim_large_solve0", "three_dim_large_solve1", "three_dim_large_solve2"]


@pytest.fixture(scope="session")
def generated_fixtures(tmp_path_factory):
    spec = importlib.util.spec_from_file_location(
        "gen_fixtures", os.path.join(GOLDEN, "gen_fixtures.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    out = str(tmp_path_factory.mktemp("generated_fixtures"))
    gen.generate(out)
    return out
