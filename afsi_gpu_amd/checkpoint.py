"""Checkpoints of the MAC/IB state and of resident serving instances.

Container and integrity scheme of mac/checkpoint.py:17-55: one compressed
``.npz`` holding a ``metadata`` JSON string and named arrays, written to
``<path>.tmp`` and atomically renamed; ``metadata['sha256']`` is the SHA-256
of the sorted-key JSON of the metadata (without the digest) followed, for
every array name in sorted order, by the name's UTF-8 bytes and the array's
C-order bytes. Loading uses ``numpy.load(allow_pickle=False)``, recomputes
the digest and checks dtypes, finiteness and the step clock as
mac/checkpoint.py:58-99 does.

``save_state``/``load_state`` carry the MAC state of mac/coupling.py:8-19
(``MACState``). ``load_state`` also reads the MAC state out of a reference
MAC checkpoint (producer 'afsi-torch-mac'); the solid-model arrays and
metadata such a file carries are verified by the digest and returned
untouched in ``extra``. ``save_resident``/``load_resident`` snapshot the
server's resident instances (grid, points, forces, velocity, pressure).
"""
from dataclasses import dataclass
import hashlib
import json
from math import isfinite
from pathlib import Path

import numpy as np
import torch

PRODUCER = 'afsi-gpu-amd-mac'
REFERENCE_PRODUCER = 'afsi-torch-mac'
SERVING_PRODUCER = 'afsi-gpu-amd-serving'
SCHEMA = 1
UNITS = 'cm-g-s'
INTEGER_ARRAYS = ('cells', 'faces', 'facet_tags')


@dataclass(frozen=True)
class MACState:
    """mac/coupling.py:8-19."""
    step: int
    time: float
    x: torch.Tensor
    velocity: tuple
    pressure: torch.Tensor
    force: torch.Tensor
    force_time: float | None
    previous_advection: tuple | None = None
    previous_dt: float | None = None
    pressure_time: float | None = None


def digest(metadata, data):
    """SHA-256 of the sorted-key JSON metadata, then name and bytes of every
    array in sorted name order (mac/checkpoint.py:17-22)."""
    result = hashlib.sha256(json.dumps(metadata, sort_keys=True, allow_nan=False).encode())
    for name in sorted(data):
        result.update(name.encode())
        result.update(np.ascontiguousarray(data[name]).tobytes())
    return result.hexdigest()


def _write(path, metadata, data):
    metadata = dict(metadata)
    metadata['sha256'] = digest(metadata, data)
    path = Path(path)
    temporary = path.with_suffix(path.suffix + '.tmp')
    with temporary.open('wb') as stream:
        np.savez_compressed(stream, metadata=json.dumps(metadata, allow_nan=False), **data)
    temporary.replace(path)


def _read(path, producers):
    with np.load(path, allow_pickle=False) as archive:
        data = {name: archive[name] for name in archive.files}
    if 'metadata' not in data:
        raise ValueError('invalid MAC checkpoint or checksum')
    metadata = json.loads(str(data.pop('metadata')))
    checksum = metadata.pop('sha256', None)
    if (metadata.get('producer') not in producers or metadata.get('schema') != SCHEMA or
            metadata.get('units') != UNITS or digest(metadata, data) != checksum):
        raise ValueError('invalid MAC checkpoint or checksum')
    for name, value in data.items():
        dtype = np.int64 if name in INTEGER_ARRAYS else np.float64
        if value.dtype != dtype or not np.isfinite(value).all():
            raise ValueError('invalid MAC checkpoint array')
    return metadata, data


def _array(t):
    return t.detach().cpu().numpy()


def save_state(path, state, settings, progress, *, extra=None):
    """Write ``state`` with ``settings`` (must hold ``dt``) and ``progress``;
    ``extra`` = further named FP64/int64 arrays to carry (e.g. geometry)."""
    check_settings(settings)
    data = {name: _array(getattr(state, name)) for name in ('x', 'pressure', 'force')}
    data.update({f'velocity_{c}': _array(u) for c, u in enumerate(state.velocity)})
    if state.previous_advection is not None:
        data.update({f'previous_advection_{c}': _array(a) for c, a in enumerate(state.previous_advection)})
    for name, value in (extra or {}).items():
        if name in data or name == 'metadata':
            raise ValueError(f'extra array {name!r} collides with the state')
        data[name] = np.asarray(value)
    metadata = dict(producer=PRODUCER, schema=SCHEMA, units=UNITS, step=state.step, time=state.time,
                    force_time=state.force_time, previous_dt=state.previous_dt,
                    pressure_time=state.pressure_time, settings=settings, progress=progress)
    _write(path, metadata, data)


def load_state(path, device='cpu'):
    """(state, settings, progress, extra): the MAC state of a checkpoint
    written by ``save_state`` or by the reference (mac/checkpoint.py:58-99
    clock rules; ``extra`` holds every other array and, for reference files,
    the solid-model metadata)."""
    metadata, data = _read(path, (PRODUCER, REFERENCE_PRODUCER))
    tensor = lambda name: torch.as_tensor(data.pop(name), device=device)
    state = MACState(metadata['step'], metadata['time'], tensor('x'),
                     tuple(tensor(f'velocity_{c}') for c in range(3)), tensor('pressure'), tensor('force'),
                     metadata['force_time'],
                     tuple(tensor(f'previous_advection_{c}') for c in range(3))
                     if 'previous_advection_0' in data else None,
                     metadata.get('previous_dt'), metadata.get('pressure_time'))
    settings = metadata['settings']
    dt = settings['dt']
    cnab = settings.get('coupling', {}).get('scheme', 'explicit-lagged') == 'cnab-midpoint'
    if cnab and state.step > 0:
        if state.previous_advection is None or state.previous_dt is None or state.pressure_time is None:
            raise ValueError('CNAB checkpoint is missing multistep history')
        for c, a in enumerate(state.previous_advection):
            if a.shape != state.velocity[c].shape:
                raise ValueError('invalid CNAB advection history shape')
        if abs(state.previous_dt - dt) > 1e-12 or abs(state.pressure_time - (state.time - .5 * dt)) > 1e-12:
            raise ValueError('invalid CNAB history clock')
    lagged = None if state.force_time is None else (state.time if cnab else (state.step - 1) * dt)
    if (type(state.step) is not int or state.step < 0 or abs(state.time - state.step * dt) > 1e-12 or
            (state.step == 0 and state.force_time is not None) or
            (state.step > 0 and (state.force_time is None or abs(state.force_time - lagged) > 1e-12))):
        raise ValueError('invalid MAC checkpoint clock')
    if state.x.ndim != 2 or state.x.shape[1] != 3 or state.force.shape != state.x.shape:
        raise ValueError('invalid MAC checkpoint array')
    extra = dict(data)
    if metadata['producer'] == REFERENCE_PRODUCER:
        extra['reference_metadata'] = {k: v for k, v in metadata.items() if k not in ('settings', 'progress')}
    return state, settings, metadata['progress'], extra


def save_resident(path, instances, settings):
    """Snapshot of resident serving instances: ``instances`` maps an integer
    id to a dict with ``kind`` ('3d' or '2d'), ``shape``, ``x``, ``force``,
    ``velocity`` (face arrays) and ``pressure`` (None before the first
    solve)."""
    data, index = {}, {}
    for key, inst in instances.items():
        tag = f'i{int(key)}'
        data[f'{tag}_x'] = _array(inst['x'])
        data[f'{tag}_force'] = _array(inst['force'])
        for c, u in enumerate(inst['velocity']):
            data[f'{tag}_velocity_{c}'] = _array(u)
        if inst['pressure'] is not None:
            data[f'{tag}_pressure'] = _array(inst['pressure'])
        index[str(int(key))] = dict(kind=inst['kind'], shape=[int(n) for n in inst['shape']],
                                    components=len(inst['velocity']), pressure=inst['pressure'] is not None)
    metadata = dict(producer=SERVING_PRODUCER, schema=SCHEMA, units=UNITS, settings=settings, instances=index)
    check_settings(settings)
    _write(path, metadata, data)


def load_resident(path, device='cpu'):
    """(instances, settings) as given to ``save_resident``."""
    metadata, data = _read(path, (SERVING_PRODUCER,))
    tensor = lambda name: torch.as_tensor(data[name], device=device)
    instances = {}
    for key, info in metadata['instances'].items():
        tag = f'i{key}'
        if info['kind'] not in ('3d', '2d') or info['components'] != len(info['shape']):
            raise ValueError('invalid serving checkpoint index')
        instances[int(key)] = dict(kind=info['kind'], shape=tuple(info['shape']), x=tensor(f'{tag}_x'),
                                   force=tensor(f'{tag}_force'),
                                   velocity=tuple(tensor(f'{tag}_velocity_{c}') for c in range(info['components'])),
                                   pressure=tensor(f'{tag}_pressure') if info['pressure'] else None)
    return instances, metadata['settings']


def check_settings(settings):
    dt = settings.get('dt')
    if not isinstance(dt, (int, float)) or not isfinite(dt) or dt <= 0:
        raise ValueError('checkpoint settings must record a positive dt')
