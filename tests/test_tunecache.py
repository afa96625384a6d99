"""The on-disk cache of tuned launch shapes (host logic, no GPU)."""

from __future__ import annotations

import json

from pygrgl_spmv.backends.hip import tunecache


def test_context_round_trip(tmp_path):
    fields = {"artifact": "abc", "dtype": 0, "library": [1, 2]}
    ctx = tunecache.TuneContext(tmp_path, fields)
    key = tunecache.launch_key("down", 3, 16, [16, 16, 64])
    assert ctx.get(key) is None
    ctx.put(key, (4, 4, 16, 8, 1, 128 | (6 << 8)))
    ctx.flush()
    again = tunecache.TuneContext(tmp_path, dict(fields))
    assert again.path == ctx.path and again.get(key) == (4, 4, 16, 8, 1, 128 | (6 << 8))


def test_other_context_does_not_share_shapes(tmp_path):
    a = tunecache.TuneContext(tmp_path, {"artifact": "abc", "dtype": 0})
    a.put(tunecache.launch_key("up", 1, 8, [8]), (2, 4, 8, 8, 1, 0))
    a.flush()
    b = tunecache.TuneContext(tmp_path, {"artifact": "abc", "dtype": 1})
    assert b.path != a.path and b.get(tunecache.launch_key("up", 1, 8, [8])) is None


def test_foreign_file_content_is_ignored(tmp_path):
    fields = {"artifact": "abc"}
    ctx = tunecache.TuneContext(tmp_path, fields)
    ctx.path.write_text(json.dumps({"context": {"artifact": "other"}, "shapes": {"x": [1, 2, 3, 4, 5, 6]}}))
    assert tunecache.TuneContext(tmp_path, fields).get("x") is None
    ctx.path.write_text("not json")
    assert tunecache.TuneContext(tmp_path, fields).get("x") is None


def test_unwritable_directory_keeps_memory_copy(tmp_path):
    blocked = tmp_path / "file"
    blocked.write_text("")
    ctx = tunecache.TuneContext(blocked / "sub", {"artifact": "abc"})
    ctx.put("k", (1, 1, 1, 8, 1, 0))
    ctx.flush()  # cannot create the directory: no exception, the shape stays in memory
    assert ctx.get("k") == (1, 1, 1, 8, 1, 0)


def test_file_digest_depends_on_content(tmp_path):
    p, q = tmp_path / "a", tmp_path / "b"
    p.write_bytes(b"x" * 1000)
    q.write_bytes(b"x" * 999 + b"y")
    assert tunecache.file_digest(p) != tunecache.file_digest(q)
    assert tunecache.file_digest(p) == tunecache.file_digest(p)


def test_default_directory_follows_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("GRG_HIP_TUNE_CACHE", str(tmp_path))
    assert tunecache.default_directory() == tmp_path
    monkeypatch.setenv("GRG_HIP_TUNE_CACHE", "")
    assert tunecache.default_directory() is None
