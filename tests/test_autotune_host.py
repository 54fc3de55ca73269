"""Host side of the autotuner: the descent's candidates, the workload key and the
JSON cache (no device calls)."""
from types import SimpleNamespace

from mslr_hip import autotune as T


def _op(N=(90, 90, 60), Nt=300, Nc=18, K=60750, dev="MI355X"):
    ctx = SimpleNamespace(info=lambda: {"name": dev})
    return SimpleNamespace(N=N, Nt=Nt, Nc=Nc, K=K, ctx=ctx)


def test_descent_tries_every_other_value_once():
    order = T.descent_order()
    assert len(order) == sum(len(v) - 1 for v in T.SPACE.values())
    assert len(set(order)) == len(order)
    for name, value in order:
        assert value != T.BUILTIN[name] and value in T.SPACE[name]
    # gram chunks are multiples of 256 in [256, 65536], as the library requires
    for v in T.SPACE["gram_chunk"]:
        assert v == 0 or (v % 256 == 0 and 256 <= v <= 65536)


def test_workload_key_separates_shapes_and_forms():
    ps, ss = [[90, 90, 60], [10, 10, 10], [1, 1, 1]], [[90, 90, 60], [10, 10, 10], [1, 1, 1]]
    k = T.workload_key(_op(), ps, ss, "mixed")
    assert k == T.workload_key(_op(), ps, ss, "mixed")
    assert k != T.workload_key(_op(), ps, ss, "batched")
    assert k != T.workload_key(_op(Nt=387), ps, ss, "mixed")
    assert k != T.workload_key(_op(dev="other"), ps, ss, "mixed")
    assert k != T.workload_key(_op(), ps[:2], ss[:2], "mixed")


def test_cache_roundtrip(tmp_path):
    path = tmp_path / "t.json"
    assert T.load_cache(path) == {}
    assert T.load_cache(None) == {}
    e1 = {"tuning": dict(T.BUILTIN, gram_chunk=8192), "ms_per_iter": 47.0}
    e2 = {"tuning": dict(T.BUILTIN), "ms_per_iter": 12.0}
    T.store_cache(path, "a", e1)
    T.store_cache(path, "b", e2)
    T.store_cache(None, "c", e2)
    c = T.load_cache(path)
    assert c == {"a": e1, "b": e2}


def test_relevance_skips_shapes_that_cannot_matter():
    ps1 = [[90, 90, 60], [10, 10, 10], [1, 1, 1]]
    ps2 = [[90, 90, 60], [30, 30, 20], [10, 10, 10], [6, 6, 6]]
    assert T.relevant("yz_threads_fwd", (90, 90, 60), 300, ps1)
    assert not T.relevant("yz_threads_adj", (32, 32, 24), 40, ps1)
    assert not T.relevant("sb2st_schedule", (90, 90, 60), 300, ps1)   # m <= 300 everywhere
    assert T.relevant("sb2st_schedule", (90, 90, 60), 387, ps2)       # m = 387
    assert T.relevant("gram_chunk", (12, 12, 10), 10, ps1)
    assert T.relevant("apply_fp32", (90, 90, 60), 300, ps1, "mixed")
    assert not T.relevant("apply_fp32", (90, 90, 60), 300, ps1, "plain")
