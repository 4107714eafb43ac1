"""Transport.class_table_bytes: the cap reaches mcx_set_model through the
environment variable MCX_CLASS_TABLE_BYTES for that call only (host logic;
the library is replaced by a recorder, no GPU)."""
import os

from mcmax_amd import configs, transport


class _Recorder:
    def __init__(self):
        self.seen = []

    def mcx_set_tables(self, ctx, tab):
        return 0

    def mcx_set_model(self, ctx, grid):
        self.seen.append(os.environ.get("MCX_CLASS_TABLE_BYTES"))
        return 0


def _transport(lib):
    t = object.__new__(transport.Transport)
    t.lib, t.ctx, t._keep, t.model = lib, None, [], None
    t.class_table_bytes = None
    return t


def test_cap_only_during_the_call(tmp_path, monkeypatch):
    monkeypatch.delenv("MCX_CLASS_TABLE_BYTES", raising=False)
    m = configs.load_config("tiny", str(tmp_path))
    lib = _Recorder()
    t = _transport(lib)
    t.set_model(m)
    t.class_table_bytes = 1 << 20
    t.set_model(m, tables=False)
    assert "MCX_CLASS_TABLE_BYTES" not in os.environ
    monkeypatch.setenv("MCX_CLASS_TABLE_BYTES", "12345")
    t.set_model(m, tables=False)
    assert os.environ["MCX_CLASS_TABLE_BYTES"] == "12345"
    t.class_table_bytes = None
    t.set_model(m, tables=False)
    assert lib.seen == [None, str(1 << 20), str(1 << 20), "12345"]
