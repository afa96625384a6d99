"""On-disk cache of the launch shapes the per-launch tuner chose.

The tuner (``HipRuntime._tune``) times every admissible launch shape of every
level and write-out launch when an op is prepared; on the 500k x 5M workload
that is about 16 s per prepared op. The chosen shapes depend only on the
artifact's content, the runtime's structural parameters (node order, split and
tile limits, residency of each block), the value type, the kernel library
build and the GPU, so they are stored per such *context* and reused by later
runtimes. A context is a file ``<directory>/<context digest>.json`` mapping
each launch key (sweep configuration, level, k, leading dimensions) to its
shape. A stale or foreign entry cannot produce a wrong result: every cached
shape is one of the kernel's valid forms and is re-checked for admissibility
before use; at worst it is slower than a fresh choice.

No reference counterpart: the reference re-tunes one configuration per
direction at every ``__enter__`` (``triton/backend.py:1062-1099``).
"""

from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path


def default_directory() -> Path | None:
    """``$GRG_HIP_TUNE_CACHE`` if set (empty: no cache), else ``~/.cache/pygrgl_spmv/hip_tune``."""
    env = os.environ.get("GRG_HIP_TUNE_CACHE")
    if env is not None:
        return Path(env) if env else None
    return Path.home() / ".cache" / "pygrgl_spmv" / "hip_tune"


def file_digest(path) -> str:
    """Content digest of a file (xxh64 when xxhash is importable, else BLAKE2b), streamed in
    64 MiB pieces."""
    try:
        import xxhash

        h = xxhash.xxh64()
    except ImportError:  # the digest only has to be stable, not a particular function
        h = hashlib.blake2b(digest_size=8)
    with open(path, "rb") as fh:
        while True:
            chunk = fh.read(64 << 20)
            if not chunk:
                break
            h.update(chunk)
    return h.hexdigest()


def launch_key(cfg: str, level: int, k: int, lds) -> str:
    return json.dumps([str(cfg), int(level), int(k), [int(v) for v in lds]])


class TuneContext:
    """The cached shapes of one context, loaded once and written back when new ones arrive."""

    def __init__(self, directory: Path, fields: dict) -> None:
        blob = json.dumps(fields, sort_keys=True).encode()
        self.digest = hashlib.blake2b(blob, digest_size=16).hexdigest()
        self.path = Path(directory) / f"{self.digest}.json"
        self.fields = fields
        self.shapes: dict[str, list[int]] = {}
        self._dirty = False
        try:
            with open(self.path) as fh:
                data = json.load(fh)
            if data.get("context") == fields:
                self.shapes = {k: list(v) for k, v in data.get("shapes", {}).items()}
        except (OSError, ValueError):
            self.shapes = {}

    def get(self, key: str):
        v = self.shapes.get(key)
        return None if v is None else tuple(int(x) for x in v)

    def put(self, key: str, shape) -> None:
        self.shapes[key] = [int(x) for x in shape]
        self._dirty = True

    def flush(self) -> None:
        """Write the context file (atomically: a sibling file renamed over it). Failures to
        write (read-only home, full disk) leave the cache in memory only."""
        if not self._dirty:
            return
        tmp = self.path.with_name(f".{self.path.name}.{os.getpid()}.tmp")
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            with open(tmp, "w") as fh:
                json.dump({"context": self.fields, "shapes": self.shapes}, fh)
            os.replace(tmp, self.path)
            self._dirty = False
        except OSError:
            try:
                tmp.unlink()
            except OSError:
                pass


__all__ = ["TuneContext", "default_directory", "file_digest", "launch_key"]
