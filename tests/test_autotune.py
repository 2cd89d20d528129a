"""The autotuner (nls_amd/autotune.py): candidate sets, choice of the fastest,
the on-disk cache.  CPU tests inject the timer; the GPU test times the real
engines and checks that the tuned solve still matches the oracle."""
import json

import numpy as np
import pytest

from nls_amd import autotune, configs, native


def fake_timer(table, calls):
    def t(cand, _setup):
        calls.append(cand)
        return table[cand[1]]
    return t


def test_plain_candidates_follow_what_fits():
    T = autotune.Tuner(path=None, device_name="dev")
    c = configs.model_coeffs(dynamic=True)
    assert T.plain_candidates(native.ProblemDesc.make(2, 64, 5, 1e-3, 0.1, c)) == ["resident", "fused", "stage"]
    assert T.plain_candidates(native.ProblemDesc.make(2, 128, 5, 1e-3, 0.1, c)) == ["fused", "stage"]
    assert T.plain_candidates(native.ProblemDesc.make(2, 64, 7, 1e-3, 0.1, c, dtype="f64")) == ["fused", "stage"]
    assert T.plain_candidates(native.ProblemDesc.make(1, 4096, 5, 1e-3, 0.1, c)) == ["resident", "stage"]
    assert T.plain_candidates(native.ProblemDesc.make(1, 5000, 5, 1e-3, 0.1, c)) == ["stage"]


def test_fastest_candidate_is_chosen_and_cached(tmp_path):
    path = str(tmp_path / "tune.json")
    calls = []
    T = autotune.Tuner(path=path, reps=2, timer=fake_timer({"resident": 3e-6, "fused": 1e-6, "stage": 2e-6}, calls),
                       device_name="dev")
    d = native.ProblemDesc.make(2, 64, 5, 1e-3, 0.1, configs.model_coeffs(dynamic=True))
    assert T.plain_engine(d) == "fused"
    assert len(calls) == 6                       # 3 candidates x 2 repetitions
    assert T.tuned(d).engine == native.ENGINE["fused"]
    # a new tuner reads the choice from disk and measures nothing
    calls2 = []
    T2 = autotune.Tuner(path=path, timer=fake_timer({"resident": 0, "fused": 1, "stage": 1}, calls2), device_name="dev")
    assert T2.plain_engine(d) == "fused" and calls2 == []
    # another device or problem class is measured anew
    T3 = autotune.Tuner(path=path, timer=fake_timer({"resident": 0, "fused": 1, "stage": 1}, calls2), device_name="other")
    assert T3.plain_engine(d) == "resident" and len(calls2) == 9
    table = json.load(open(path))
    assert len(table) == 2 and all("seconds_per_step" in v for v in table.values())


def test_stage_tile_and_batch_choices():
    calls = []
    T = autotune.Tuner(path=None, timer=lambda cand, _s: {4: 4.0, 8: 3.0, 16: 1.0, 32: 2.0}[cand[1]], device_name="dev")
    assert T.stage_tile("f32") == 16
    ms = configs.sweep512(count=4, steps=10)
    T = autotune.Tuner(path=None, timer=fake_timer({"resident": 9, "fused": 1, "stage": 2, "hybrid": 3}, calls),
                       device_name="dev")
    assert T.batch_engine(ms) == "fused"
    assert all(c[1] != "resident" for c in calls) or all(m.n == 64 for m in ms)
    # a mix of members that fit the resident engine and members that do not
    # adds the hybrid engine to the candidates; all fit: resident, no hybrid
    mixed = [m for m in configs.sweep512(count=64, steps=10)]
    assert any(m.n <= 64 for m in mixed) and any(m.n > 64 for m in mixed)
    calls.clear()
    T = autotune.Tuner(path=None, timer=fake_timer({"resident": 9, "fused": 2, "stage": 3, "hybrid": 1}, calls),
                       device_name="dev")
    assert T.batch_engine(mixed) == "hybrid"
    assert sorted({c[1] for c in calls if c[0] == "batch"}) == ["fused", "hybrid", "stage"]
    small = [m for m in mixed if m.n <= 64]
    calls.clear()
    assert T.batch_engine(small) == "fused"               # fused 2 < stage 3 < resident 9
    assert sorted({c[1] for c in calls if c[0] == "batch"}) == ["fused", "resident", "stage"]


def test_fused_form_candidates_and_choice():
    T = autotune.Tuner(path=None, device_name="dev",
                       timer=lambda cand, _s: {"tall": 2.0, "short": 3.0, "half": 1.0, "tall512": 2.5}[cand[1]])
    assert T.form_candidates("f32", "periodic", 5) == ["tall", "short", "tall512"]
    assert T.form_candidates("f32", "periodic", 3) == ["tall", "short"]
    assert T.form_candidates("f32", "ref", 5) == ["tall", "short"]
    assert T.form_candidates("f64", "periodic", 5) == ["tall", "short", "half"]
    assert T.form_candidates("f64", "ref", 3) == ["tall", "short"]
    assert T.form_candidates("f64", "periodic", 7) == ["tall"]
    ms = configs.sweep512(count=4, steps=10)
    assert T.fused_form(ms, dtype="f32") == "tall"
    assert T.fused_form(ms, dtype="f64") == "half"
    d = native.ProblemDesc.make(2, 2048, 5, 1e-4, 0.05, configs.model_coeffs(), bc="periodic", dtype="f64")
    assert T.plain_fused_form(d) == "half"


@pytest.mark.gpu
def test_tuned_fused_form_batch_matches_each_form(gpu):
    """The autotuner times every tile form of the fused kernel on a batch,
    applies the fastest, and a batch created then runs that form; results are
    the same in every form, bit for bit."""
    from nls_amd.batch import Batch
    ms = configs.sweep512(count=24, steps=40)
    T = autotune.Tuner(path=None, steps=16, reps=1)
    try:
        best = T.fused_form(ms, dtype="f32", engine="fused")
        assert best in ("tall", "short", "tall512") and set(T.last) == {"tall", "short", "tall512"}
        assert native.fused_form("f32") == best
        out = {}
        for form, ty in (("tall", 32), ("short", 16), ("tall512", 32)):
            native.fused_form("f32", form)
            b = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine="fused")
            assert b.tiles == sum(((m.n + 63) // 64) * ((m.n + ty - 1) // ty) for m in ms)
            b.advance(40)
            out[form] = b.psi.cpu().numpy()
            b.close()
        np.testing.assert_array_equal(out["tall"], out["short"])
        np.testing.assert_array_equal(out["tall"], out["tall512"])
    finally:
        native.fused_form("f32", "auto")


@pytest.mark.gpu
@pytest.mark.parametrize("n", [40, 72, 130, 256])
def test_plain_tall512_form_equals_tall(gpu, n):
    """The 512-thread form of the single-precision periodic pair kernel gives
    the 1024-thread form's result bit for bit on a plain solve (grids smaller
    than one tile, not a multiple of it, several tiles); in double precision
    and with reflecting boundaries it runs as the tall form."""
    X, Y, h = configs.grid_2d(n, 10.0)
    r2 = X ** 2 + Y ** 2
    rng = np.random.default_rng(n)
    psi0 = 0.3 * np.exp(-r2 / 4.0) + configs.noise(rng, (n, n), 1e-2)
    P = 2.0 * np.exp(-r2 / 6.0)
    n0 = 0.2 * np.exp(-r2 / 3.0)
    try:
        for dtype, bc, dyn in (("f32", "periodic", True), ("f32", "periodic", False), ("f64", "periodic", True),
                               ("f32", "ref", True)):
            c = configs.model_coeffs(dynamic=dyn)
            out = {}
            for form in ("tall", "tall512"):
                native.fused_form(dtype, form)
                d = native.ProblemDesc.make(2, n, 5, 2e-4, h, c, bc=bc, dtype=dtype, engine="fused")
                out[form] = native.solve_host(d, 37, P, psi0, n0)
            for a, b in zip(out["tall"], out["tall512"]):
                if a is None:                      # no reservoir field without the reservoir equation
                    assert b is None
                    continue
                assert np.all(np.isfinite(a))
                np.testing.assert_array_equal(a, b)
    finally:
        native.fused_form("f32", "auto")
        native.fused_form("f64", "auto")


@pytest.mark.gpu
def test_tuned_plain_solve_matches_oracle(gpu, oracle, tmp_path):
    T = autotune.Tuner(path=str(tmp_path / "t.json"), steps=64, reps=1)
    n = 48
    X, Y, h = configs.grid_2d(n, 10.0)
    r2 = X ** 2 + Y ** 2
    rng = np.random.default_rng(3)
    psi0 = 0.3 * np.exp(-r2 / 4.0) + configs.noise(rng, (n, n), 1e-2)
    P = 2.0 * np.exp(-r2 / 6.0)
    n0 = 0.2 * np.exp(-r2 / 3.0)
    c = configs.model_coeffs(dynamic=True)
    d = native.ProblemDesc.make(2, n, 5, 1e-3, h, c, bc="periodic", dtype="f64")
    engine = T.plain_engine(d)
    assert engine in ("resident", "fused", "stage") and set(T.last) == {"resident", "fused", "stage"}
    psi, nres = native.solve_host(T.tuned(d), 200, P, psi0, n0)
    opsi, _ = oracle.rk4(2, n, 5, 1, 1e-3, h, 200, P, c, psi0, n0)
    assert np.linalg.norm(psi - opsi) / np.linalg.norm(opsi) <= 1e-10
    rows = T.stage_tile("f64", n=256)
    assert rows in (4, 8, 16, 32) and native.stage_tile("f64") == rows
    native.stage_tile("f64", 16)
