"""The request loop's admission (breseq_amd.batch.Server) on the host, with the sample pipeline
replaced by a stub: never more than `parallel` requests running, never more estimated device memory
in use than the budget, every request finished, failures reported per request."""
import threading
import time

from breseq_amd import batch


def test_server_admission(monkeypatch):
    sizes = {i: (3 + (i % 4)) << 20 for i in range(12)}
    state = {"running": 0, "bytes": 0, "max_running": 0, "max_bytes": 0}
    lock = threading.Lock()

    def fake_bytes(sample):
        return sizes[sample]

    def fake_pipeline(pool, sample, mode, dp):
        with lock:
            state["running"] += 1
            state["bytes"] += sizes[sample]
            state["max_running"] = max(state["max_running"], state["running"])
            state["max_bytes"] = max(state["max_bytes"], state["bytes"])
        time.sleep(0.02)
        with lock:
            state["running"] -= 1
            state["bytes"] -= sizes[sample]
        if sample == 5:
            raise RuntimeError("failing request")
        return {"gd": "gd%d" % sample, "dp": None}

    monkeypatch.setattr(batch, "device_bytes", fake_bytes)
    monkeypatch.setattr(batch, "run_pooled", fake_pipeline)
    srv = batch.Server(parallel=3, hbm_budget=12 << 20)
    for i in range(12):
        srv.submit(i, i)
    results, errors = srv.close()
    assert set(results) == set(range(12)) - {5}
    assert set(errors) == {5}
    assert all(results[i]["gd"] == "gd%d" % i and results[i]["latency_s"] > 0 for i in results)
    assert state["max_running"] <= 3
    assert state["max_bytes"] <= 12 << 20


def test_server_refuses_oversized(monkeypatch):
    monkeypatch.setattr(batch, "device_bytes", lambda s: 100)
    srv = batch.Server(parallel=1, hbm_budget=50)
    try:
        srv.submit(0, object())
        assert False, "an oversized request must be refused"
    except ValueError:
        pass
    srv.close()
