import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PARENT = os.path.join(REPO, "aerial-3d-model_amd")
for p in (PKG_PARENT, os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


_SURVEYS = {}


def survey_cached(name):
    """Render a named synthetic survey once per session."""
    from aerial_mvs import survey
    if name not in _SURVEYS:
        _SURVEYS[name] = survey.make_survey(name)
    return _SURVEYS[name]


@pytest.fixture(scope="session")
def tiny_survey():
    return survey_cached("tiny")


@pytest.fixture(scope="session")
def small_survey():
    return survey_cached("small")


@pytest.fixture(scope="session")
def gpu_ctx():
    from aerial_mvs import capi
    ctx = capi.Context(0)
    yield ctx
    ctx.close()
