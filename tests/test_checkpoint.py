"""Checkpoint container and integrity scheme (mac/checkpoint.py:17-99): the
digest against an explicit SHA-256 of the byte stream, round trips of the
MAC state and of serving snapshots, reading the MAC state out of a
reference-format file, and the reference's rejection paths (checksum,
dtype, non-finite, clock, CNAB history)."""
import hashlib
import json

import numpy as np
import pytest
import torch

from afsi_gpu_amd import checkpoint as ck


def state(step=2, dt=1e-3, cnab=False, seed=0):
    g = np.random.default_rng(seed)
    shapes = ((9, 8, 8), (8, 9, 8), (8, 8, 9))
    velocity = tuple(torch.tensor(g.normal(size=s)) for s in shapes)
    history = tuple(torch.tensor(g.normal(size=s)) for s in shapes) if cnab else None
    force_time = None if step == 0 else (step * dt if cnab else (step - 1) * dt)
    return ck.MACState(step, step * dt, torch.tensor(g.normal(size=(50, 3))), velocity,
                       torch.tensor(g.normal(size=(8, 8, 8))), torch.tensor(g.normal(size=(50, 3))), force_time,
                       history, dt if cnab else None, step * dt - .5 * dt if cnab else None)


def test_digest_is_sha256_of_sorted_json_then_sorted_named_arrays():
    metadata = {'b': 1, 'a': [1.5, 'x']}
    data = {'zeta': np.arange(3, dtype=np.float64), 'alpha': np.array([[1, 2], [3, 4]], dtype=np.int64)}
    stream = b'{"a": [1.5, "x"], "b": 1}' + b'alpha' + data['alpha'].tobytes() + b'zeta' + data['zeta'].tobytes()
    assert ck.digest(metadata, data) == hashlib.sha256(stream).hexdigest()
    # order of insertion does not matter, any byte does
    assert ck.digest(dict(reversed(metadata.items())), dict(reversed(data.items()))) == ck.digest(metadata, data)
    changed = dict(data, zeta=np.array([0., 1., 2.000000001]))
    assert ck.digest(metadata, changed) != ck.digest(metadata, data)
    # C-order bytes: a transposed view hashes as its contiguous copy
    t = {'alpha': data['alpha'].T}
    assert ck.digest({}, t) == ck.digest({}, {'alpha': np.ascontiguousarray(data['alpha'].T)})


@pytest.mark.parametrize('cnab', [False, True])
def test_state_round_trip_and_atomic_write(tmp_path, cnab):
    s = state(cnab=cnab)
    settings = dict(dt=1e-3, coupling=dict(scheme='cnab-midpoint' if cnab else 'explicit-lagged'))
    path = tmp_path / 'state.npz'
    ck.save_state(path, s, settings, dict(accepted=2), extra={'cells': np.arange(8, dtype=np.int64).reshape(2, 4)})
    assert not (tmp_path / 'state.npz.tmp').exists()
    r, rs, progress, extra = ck.load_state(path)
    assert (r.step, r.time, r.force_time) == (s.step, s.time, s.force_time)
    assert rs == settings and progress == dict(accepted=2)
    for a, b in [(r.x, s.x), (r.pressure, s.pressure), (r.force, s.force), *zip(r.velocity, s.velocity)]:
        assert torch.equal(a, b)
    assert (r.previous_advection is None) == (not cnab)
    np.testing.assert_array_equal(extra['cells'], np.arange(8).reshape(2, 4))


def rewrite(path, edit):
    with np.load(path, allow_pickle=False) as archive:
        data = {k: archive[k] for k in archive.files}
    edit(data)
    np.savez_compressed(path, **data)


def test_rejections(tmp_path):
    path = tmp_path / 'state.npz'
    ck.save_state(path, state(), dict(dt=1e-3), {})
    rewrite(path, lambda d: d.__setitem__('x', d['x'] + 1e-15))
    with pytest.raises(ValueError, match='checksum'):
        ck.load_state(path)
    # a consistent digest over a wrong dtype / non-finite value is still rejected
    for bad in (lambda d: d['force'].astype(np.float32), lambda d: np.where(d['force'] > 0, np.nan, d['force'])):
        ck.save_state(path, state(), dict(dt=1e-3), {})

        def edit(d):
            meta = json.loads(str(d['metadata']))
            meta.pop('sha256')
            d['force'] = bad(d)
            arrays = {k: v for k, v in d.items() if k != 'metadata'}
            meta['sha256'] = ck.digest(meta, arrays)
            d['metadata'] = np.array(json.dumps(meta))
        rewrite(path, edit)
        with pytest.raises(ValueError, match='invalid MAC checkpoint array'):
            ck.load_state(path)
    s = state()
    wrong = ck.MACState(s.step, s.time, s.x, s.velocity, s.pressure, s.force, s.force_time + 1e-3)
    ck.save_state(path, wrong, dict(dt=1e-3), {})
    with pytest.raises(ValueError, match='clock'):
        ck.load_state(path)
    s = state(cnab=True)
    ck.save_state(path, ck.MACState(*[getattr(s, f) for f in ('step', 'time', 'x', 'velocity', 'pressure', 'force',
                                                             'force_time')]),
                  dict(dt=1e-3, coupling=dict(scheme='cnab-midpoint')), {})
    with pytest.raises(ValueError, match='multistep history'):
        ck.load_state(path)
    with pytest.raises(ValueError, match='positive dt'):
        ck.save_state(path, state(), dict(dt=0.), {})


def test_reads_mac_state_of_a_reference_format_checkpoint(tmp_path):
    """A file laid out as mac/checkpoint.py:25-55 writes it (solid arrays,
    producer 'afsi-torch-mac'), built here with numpy, json and hashlib."""
    s = state(step=3)
    g = np.random.default_rng(5)
    data = {'X': g.normal(size=(12, 3)), 'cells': np.arange(40, dtype=np.int64).reshape(4, 10) % 12,
            'faces': np.zeros((2, 6), dtype=np.int64), 'facet_tags': np.ones(2, dtype=np.int64),
            'fiber': g.normal(size=(12, 3)), 'x': s.x.numpy(), 'pressure': s.pressure.numpy(),
            'force': s.force.numpy()}
    data.update({f'velocity_{c}': u.numpy() for c, u in enumerate(s.velocity)})
    metadata = dict(producer='afsi-torch-mac', schema=1, units='cm-g-s', geometry={'mesh_size': .5},
                    vertex_count=12, gmsh='4.11', step=3, time=s.time, force_time=s.force_time,
                    previous_dt=None, pressure_time=None, settings=dict(dt=1e-3), progress=dict(steps=3))
    h = hashlib.sha256(json.dumps(metadata, sort_keys=True, allow_nan=False).encode())
    for name in sorted(data):
        h.update(name.encode())
        h.update(data[name].tobytes())
    metadata['sha256'] = h.hexdigest()
    path = tmp_path / 'reference.npz'
    np.savez_compressed(path, metadata=json.dumps(metadata), **data)
    r, settings, progress, extra = ck.load_state(path)
    assert r.step == 3 and torch.equal(r.x, s.x) and torch.equal(r.velocity[2], s.velocity[2])
    assert settings == dict(dt=1e-3) and progress == dict(steps=3)
    assert extra['reference_metadata']['gmsh'] == '4.11'
    np.testing.assert_array_equal(extra['cells'], data['cells'])


def test_resident_snapshot_round_trip(tmp_path):
    g = np.random.default_rng(2)
    instances = {
        4: dict(kind='3d', shape=(8, 8, 8), x=torch.tensor(g.random((30, 3))), force=torch.tensor(g.normal(size=(30, 3))),
                velocity=tuple(torch.tensor(g.normal(size=(8, 8, 8))) for _ in range(3)),
                pressure=torch.tensor(g.normal(size=(8, 8, 8)))),
        9: dict(kind='2d', shape=(16, 4), x=torch.tensor(g.random((10, 2))), force=torch.tensor(g.normal(size=(10, 2))),
                velocity=tuple(torch.tensor(g.normal(size=(16, 4))) for _ in range(2)), pressure=None)}
    path = tmp_path / 'serving.npz'
    ck.save_resident(path, instances, dict(dt=1e-3))
    back, settings = ck.load_resident(path)
    assert settings == dict(dt=1e-3) and sorted(back) == [4, 9]
    for key, inst in instances.items():
        got = back[key]
        assert got['kind'] == inst['kind'] and got['shape'] == inst['shape']
        assert torch.equal(got['x'], inst['x']) and torch.equal(got['force'], inst['force'])
        assert all(torch.equal(a, b) for a, b in zip(got['velocity'], inst['velocity']))
        assert (got['pressure'] is None) == (inst['pressure'] is None)
    with pytest.raises(ValueError, match='checksum'):
        ck.load_state(path)
