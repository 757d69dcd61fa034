"""Small autotuner for the batch shape of the persistent 2D sweep.

The persistent sweep stages each slice's coefficients once per workgroup and
shares them between the workgroup's probes (``probes_per_block``): more probes
per workgroup cut coefficient traffic and barriers per probe, fewer give more
workgroups to spread over the 256 CUs and more wavefronts per SIMD.  The best
value depends on the widths, the probe counts and the number of batches in the
launch, so it is measured: each candidate runs the real launch (outputs are
simply rewritten), in interleaved rounds after one warm-up round of all of them
(a GPU that is still raising its clocks would otherwise favour the candidates
timed last), the median per candidate decides, and the choice is
cached per workload shape and device for the process (and in a JSON file when
``THICK_PTYCHO_AMD_TUNE_FILE`` names one).
"""

from __future__ import annotations

import json
import os
from typing import Dict, Sequence, Tuple

import torch

from .forward_model import tri_engine

_CACHE: Dict[str, dict] = {}


def _key(specs: Sequence["tri_engine.TriScanSpec"]) -> str:
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    shape = [(s.fac.n, int(s.u0.shape[0]), s.fac.steps, s.fac.chunk, int(s.fac.sigma),
              s.out is not None, s.grad is not None) for s in specs]
    return json.dumps([dev, shape])


def _load_file() -> None:
    path = os.environ.get("THICK_PTYCHO_AMD_TUNE_FILE")
    if path and os.path.exists(path):
        with open(path) as f:
            _CACHE.update(json.load(f))


def _save_file() -> None:
    path = os.environ.get("THICK_PTYCHO_AMD_TUNE_FILE")
    if path:
        with open(path, "w") as f:
            json.dump(_CACHE, f, indent=1)


def candidates(specs) -> Tuple[int, ...]:
    chunk = max(s.fac.chunk for s in specs)
    if any(s.grad is not None for s in specs):
        return (1, 4)
    return (1, 2) if chunk >= 32 else (1, 2, 4, 8)


def tune_sweep(specs: Sequence["tri_engine.TriScanSpec"], reps: int = 5) -> dict:
    """{'ppb': best probes per workgroup, 'ms': {ppb: median ms}} for this launch shape."""
    if not _CACHE:
        _load_file()
    key = _key(specs)
    if key in _CACHE:
        return _CACHE[key]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cands = candidates(specs)
    for ppb in cands:  # warm-up round
        tri_engine.sweep(specs, probes_per_block=ppb)
    ts = {ppb: [] for ppb in cands}
    for _ in range(reps):
        for ppb in cands:
            e0.record()
            tri_engine.sweep(specs, probes_per_block=ppb)
            e1.record()
            torch.cuda.synchronize()
            ts[ppb].append(e0.elapsed_time(e1))
    times = {ppb: sorted(v)[len(v) // 2] for ppb, v in ts.items()}
    best = min(times, key=times.get)
    res = {"ppb": best, "ms": {str(k): v for k, v in times.items()}}
    _CACHE[key] = res
    _save_file()
    return res
