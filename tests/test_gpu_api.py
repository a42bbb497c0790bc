"""The package's call surface on the GPU: the functional API (register with a start pose, refine,
score, shift_map, rs_map, search, register_stack), the progress events of searches and climbs
(the check of the reference's tests/test_api.py:263-268), and its error behaviour.
"""
import json

import numpy as np
import pytest

import zcmir_amd
from zcmir_amd import synth

pytestmark = pytest.mark.gpu


def _corner_err(Ha, Hb, w, h):
    pts = np.array([[1, 1, 1], [w, 1, 1], [1, h, 1], [w, h, 1]], float).T
    a, b = Ha @ pts, Hb @ pts
    return float(np.abs(a[:2] / a[2] - b[:2] / b[2]).max())


@pytest.fixture(autouse=True)
def _fresh_shared_engine():
    yield
    zcmir_amd.close()


def test_search_emits_log_trail_pose_and_sweep_events(reg_factory):
    """A sweep search from a wrong pose finds a better one and reports it: the events hold the
    kinds log, trail, pose and sweep, each a dict parsed from one JSON line, the sweep event with
    the grid, the per-cell landscape and the result."""
    mv, fx, Ht = synth.rotscale_pair((192, 192), theta_deg=30.0, scale=1.05, seed=7)
    seen = []
    reg = reg_factory(half=False, sw_nth=24, sw_ns=3, sw_smin=0.9, sw_smax=1.1, sw_grid=64)
    reg.on_event = seen.append
    reg.set_images(mv, fx, preprocess=False)
    reg.pose = np.array([[1, 0, 20.0], [0, 1, -15.0], [0, 0, 1.0]])
    r = reg.search()
    assert r.improved
    ev = reg.events()
    kinds = {e["kind"] for e in ev}
    assert {"log", "trail", "pose", "sweep"} <= kinds, kinds
    assert ev == seen   # the callback saw the same events, in order
    sw = next(e for e in ev if e["kind"] == "sweep")
    assert len(sw["thetas"]) == 24 and len(sw["sigmas"]) == 3 and len(sw["best"]) == 72
    assert sw["applied"] == 0 and len(sw["result"]) == 2 and len(sw["cands"]) >= 1
    tr = [e for e in ev if e["kind"] == "trail"][-1]
    assert tr["label"] == "search" and np.allclose(np.reshape(tr["H"], (3, 3)), r.H)
    assert reg.events() == []   # taken
    json.dumps(ev)   # plain JSON values


def test_climb_emits_events(reg_factory):
    mv, fx, _ = synth.shift_pair((160, 160), (6, -4), seed=2)
    reg = reg_factory(half=False, hop=True)
    reg.set_images(mv, fx, preprocess=False)
    reg.pose = np.eye(3)
    c = reg.climb()
    ev = reg.events()
    logs = [e["text"] for e in ev if e["kind"] == "log"]
    assert any(t.startswith("ascent done") for t in logs)
    if c.moved:
        assert any(e["kind"] == "trail" for e in ev) and any(e["kind"] == "pose" for e in ev)


def test_register_from_a_pose_and_refine():
    """register(init=H) climbs from H to the true translation; refine is the same call;
    register without init names the missing keypoint stage."""
    mv, fx, Ht = synth.shift_pair((192, 192), (9, -5), seed=3)
    H0 = np.array([[1, 0, -6.0], [0, 1, 2.0], [0, 0, 1.0]])
    r = zcmir_amd.register(mv, fx, init=H0, preprocess=False, half=False)
    assert r.method == "refine" and r.iterations >= 1
    assert _corner_err(r.H, Ht, 192, 192) < 0.5
    r2 = zcmir_amd.refine(mv, fx, H0, preprocess=False, half=False)
    assert np.allclose(r2.H, r.H, atol=1e-9)
    assert abs(zcmir_amd.score(mv, fx, r.H, preprocess=False, half=False) - r.score) < 1e-9 * abs(r.score) + 1e-12
    with pytest.raises(zcmir_amd.EngineError, match="init"):
        zcmir_amd.register(mv, fx, preprocess=False)
    with pytest.raises(TypeError, match="did you mean"):
        zcmir_amd.register(mv, fx, init=H0, metrc="smi")
    # the result record's geometry: the displacement field maps fixed pixels onto the moving image
    d = r.displacement()
    assert d.shape == (192, 192, 2)
    p = r.fixed_to_moving([[50.0, 60.0]])
    assert np.allclose(p, np.array([[50.0, 60.0]]) + d[59, 49], atol=1e-4)
    w = r.warp(mv)
    assert w.shape == mv.shape and w.dtype == mv.dtype


def test_functional_maps_and_search():
    mv, fx, Ht = synth.shift_pair((160, 160), (7, -3), seed=5)
    sm = zcmir_amd.shift_map(mv, fx, preprocess=False, half=False, map_res=512)
    assert (sm.dx, sm.dy) == (7.0, -3.0)
    rs = zcmir_amd.rs_map(mv, fx, sm.corrected(np.eye(3)), preprocess=False, half=False)
    assert abs(rs.dth) < 0.05 and abs(rs.dsg) < 0.05
    mv2, fx2, Ht2 = synth.rotscale_pair((192, 192), theta_deg=-25.0, scale=1.0, seed=8)
    s = zcmir_amd.search(mv2, fx2, rotation=60, scale=(0.95, 1.05), preprocess=False, half=False, sw_nth=25,
                         sw_ns=3, sw_grid=64)
    assert s.method == "search"
    assert _corner_err(s.H, Ht2, 192, 192) < 3.0   # the reference checks its searches at 5 px
    with pytest.raises(zcmir_amd.EngineError, match="cloud"):
        zcmir_amd.search(mv2, fx2, mode="cloud", preprocess=False)


def test_register_stack_batched():
    """Four drifting slices registered together: every slice within 1 px of its pose, the
    records in slice order, the trend flags computed, the JSON export readable."""
    sl = synth.drift_stack(4, 256, seed=21)
    st = zcmir_amd.register_stack([s[0] for s in sl], [s[1] for s in sl], preprocess=False, half=False)
    assert len(st) == 4 and not any(st.flagged)
    errs = [_corner_err(r.H, Ht, 256, 256) for r, (_, _, Ht) in zip(st.slices, sl)]
    assert max(errs) < 1.0, errs
    doc = json.loads(st.to_json())
    assert len(doc["slices"]) == 4 and doc["slices"][0]["method"] in ("stack", "carry")
