import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def model_dirs(tmp_path_factory):
    """oss_input directories of the seeded synthetic models, built on demand."""
    from livespikesorter_amd import synth
    cache = {}

    def get(name, **kw):
        key = (name, tuple(sorted(kw.items())))
        if key not in cache:
            base = dict(synth.CONFIGS[name]) if name in synth.CONFIGS else {}
            base.update(kw)
            seed = base.pop("seed", synth.TEST_SEED)
            m = synth.make_model(seed=seed, **base)
            d = tmp_path_factory.mktemp(name)
            synth.write_oss_input(m, d)
            cache[key] = (m, d)
        return cache[key]
    return get
