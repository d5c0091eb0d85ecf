"""Device-solved 3-D bases through the disk cache: a miss solves on the
engine and publishes ``basis_{idx}`` + ``labels`` entries, a hit serves the
same bits without touching the engine, and the entry keys on the masks'
content and the solve settings."""
import json

import numpy as np
import pytest

from ion_gym_amd.io import basis_cache as B
from ion_gym_amd.io import fa_cache as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from ion_gym_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _two_plates(n=33, nz=17):
    a = np.zeros((n, n, nz), bool)
    b = np.zeros_like(a)
    a[8:25, 10:12, 4:13] = True
    b[8:25, 21:23, 4:13] = True
    ground = np.zeros_like(a)
    ground[:, 0, :] = ground[:, -1, :] = True
    return {1: a, 2: b, 3: ground}


@pytest.mark.parametrize("method", ["sor", "mg", "vmg"])
def test_device_bases_miss_then_hit(eng, tmp_path, method):
    from ion_gym_amd.physics import solver3d as S
    masks = _two_plates()
    print(f"device bases {method} 33x33x17, 3 electrodes", flush=True)
    bases, hit = B.solve_bases_cached(masks, method=method, tol=1e-6,
                                      root=tmp_path, engine=eng, name="pl")
    assert hit is False
    if method == "sor":
        direct = S.solve_bases(masks, tol=1e-6, engine=eng)
    elif method == "vmg":
        direct = S.solve_bases_mg(masks, tol=1e-6, engine=eng)
    else:
        direct = S._bases(masks, lambda f, v: S.solve3d_mg(
            f, v, tol=1e-6, engine=eng), 1e4, None, False)
    for i in masks:
        assert np.array_equal(bases[i], direct[i]), i
    # the hit never reaches the engine
    again, hit = B.solve_bases_cached(masks, method=method, tol=1e-6,
                                      root=tmp_path, engine=object())
    assert hit is True
    for i in masks:
        assert np.array_equal(again[i], direct[i])
        assert not again[i].flags.writeable
    (entry,) = [p for p in tmp_path.iterdir()]
    meta = json.loads((entry / "meta.json").read_text())
    assert set(meta["arrays"]) == {"basis_1", "basis_2", "basis_3", "labels"}
    assert meta["arrays"]["labels"]["dtype"] == "uint8"
    assert meta["spec"]["method"] == method
    _, labels = B.load_3d(meta["spec"], root=tmp_path)
    assert int(labels[12, 11, 8]) == 1 and int(labels[12, 22, 8]) == 2 \
        and int(labels[5, 0, 3]) == 3 and int(labels[16, 16, 8]) == 0
    assert F.verify(entry.name, root=tmp_path)
    # another tolerance, or a moved plate, is another entry
    k_tol = F.spec_key(B.device_key_dict(
        masks, method=method, mirror=(False,) * 3, v_basis=1e4, tol=1e-7,
        stencil="ghost_linear"))
    moved = dict(masks)
    moved[2] = np.roll(masks[2], 1, axis=0)
    k_moved = F.spec_key(B.device_key_dict(
        moved, method=method, mirror=(False,) * 3, v_basis=1e4, tol=1e-6,
        stencil="ghost_linear"))
    assert entry.name not in (k_tol, k_moved)
