"""The sweep harness runs end to end on the GPU (small configuration)."""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_accuracy_and_perf_sweeps(gpu, tmp_path):
    from ozaie_amd import sweep
    out = tmp_path / "acc.json"
    sweep.main(["accuracy", "--phis", "1", "--seeds", "1", "--ks", "128,256", "--modes", "accurate,fast,matched,crt",
                "--samples", "200", "--out", str(out)])
    recs = [json.loads(l) for l in open(out)]
    assert len(recs) == 8
    for r in recs:
        assert r["max_rel_err"] < 1e-6 and r["native_max_rel_err"] < 1e-6
    # resume skips everything already measured
    sweep.main(["accuracy", "--phis", "1", "--seeds", "1", "--ks", "128,256", "--modes", "accurate,fast,matched,crt",
                "--samples", "200", "--out", str(out), "--resume"])
    assert len(open(out).readlines()) == 8
    pout = tmp_path / "perf.json"
    sweep.main(["perf", "--sizes", "512", "--modes", "fast3,fast,crt", "--iters", "2", "--warmup", "1",
                "--out", str(pout)])
    recs = [json.loads(l) for l in open(pout)]
    assert len(recs) == 3 and all(r["tflops"] > 0 for r in recs)
