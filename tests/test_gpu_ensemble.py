"""GPU: the batched ensemble path against independent single-solver runs and
the oracle.  Members of different sizes, viscosities and step sizes share
batched launches (size classes on concurrent streams) and a replayed hipGraph;
each member must reproduce its own plain solve (bit-identical with the same
pass shapes, 1e-13 with autotuned ones) and the oracle."""
import numpy as np
import pytest

import ns2d_oracle as O

pytestmark = pytest.mark.gpu


def cfg_of(p):
    import ns2d
    return ns2d.Config(Nx=p.Nx, Ny=p.Ny, nu=p.nu, hyper_flag=p.hyper_flag, hyper_pow=p.hyper_pow,
                       alpha=p.alpha, drag_flag=p.drag_flag, drag_pow=p.drag_pow, filter=p.filter,
                       integrator="RK4CN")


def _members():
    specs = [(64, 64, 1e-3, 1e-3), (128, 128, 2e-3, 7e-4), (64, 64, 5e-4, 1.5e-3),
             (256, 128, 1e-3, 5e-4), (128, 128, 1e-3, 1e-3), (64, 64, 1e-3, 2e-3)]
    params, w0s, dts = [], [], []
    for i, (nx, ny, nu, dt) in enumerate(specs):
        p = O.Params(nx, ny, nu=nu, hyper_flag=0, hyper_pow=1.0, filter="23")
        params.append(p)
        w0s.append(O.ic_random_spectrum(p, seed=100 + i, k0=4.0 + i))
        dts.append(dt)
    return params, w0s, dts


def test_autotuned_ensemble_matches_plain_solves_and_oracle():
    """The autotuner may pick other pass shapes than a plain solver's
    defaults (different FFT radices, so different roundoff): members then
    agree with their plain solves to 1e-13 and with the oracle to 1e-12."""
    import ns2d
    params, w0s, dts = _members()
    nsteps = 9
    with ns2d.Ensemble([cfg_of(p) for p in params]) as ens:
        for i, w in enumerate(w0s):
            ens[i].set_state(w)
        ens.step(dts, nsteps)
        got = [ens[i].get_state() for i in range(len(ens))]
        shapes = [ens.variants(i) for i in range(len(ens))]
    for i, p in enumerate(params):
        k, r = shapes[i]
        assert 0 <= k < len(ns2d.kpass_variants(p.Nx)) and 0 <= r < len(ns2d.rpass_variants(p.Ny))
        with ns2d.Solver(cfg_of(p)) as s:
            s.set_state(w0s[i])
            s.spectral_step(dts[i], nsteps)
            plain = s.get_state()
        assert np.linalg.norm(got[i] - plain) / np.linalg.norm(plain) < 1e-13, i
        st = O.State(w0s[i], p)
        O.step(st, dts[i], nsteps)
        assert np.linalg.norm(got[i] - st.w_hat) / np.linalg.norm(st.w_hat) < 1e-12


def test_ensemble_matches_plain_solves_and_oracle(monkeypatch):
    """Without the autotuner the batched launches use the plain solver's
    shapes: each member is bit-identical to its own plain solve."""
    import ns2d
    monkeypatch.setenv("NS2D_ENS_TUNE", "0")
    params, w0s, dts = _members()
    nsteps = 7    # odd: graph periods plus a single batched step
    print("ensemble of", len(params), flush=True)
    with ns2d.Ensemble([cfg_of(p) for p in params]) as ens:
        for i, w in enumerate(w0s):
            ens[i].set_state(w)
        ens.step(dts, 3)
        ens.step(dts, nsteps - 3)
        got = [ens[i].get_state() for i in range(len(ens))]
        info = [ens[i].info() for i in range(len(ens))]
    for i, p in enumerate(params):
        assert info[i]["iters"] == nsteps
        with ns2d.Solver(cfg_of(p)) as s:
            s.set_state(w0s[i])
            s.spectral_step(dts[i], nsteps)
            plain = s.get_state()
        assert np.array_equal(got[i], plain), i
        st = O.State(w0s[i], p)
        O.step(st, dts[i], nsteps)
        assert np.linalg.norm(got[i] - st.w_hat) / np.linalg.norm(st.w_hat) < 1e-12


def test_graph_replay_matches_direct_launches():
    """hipGraph replay of step periods (default) equals direct launches."""
    import os

    import ns2d
    p = O.Params(128, 128, nu=1e-3, hyper_flag=0, hyper_pow=1.0)
    w0 = O.ic_random_spectrum(p, seed=3, k0=6.0)
    with ns2d.Solver(cfg_of(p)) as s:
        s.set_state(w0)
        for _ in range(3):
            s.spectral_step(1e-3, 5)
            s.sync()
        a = s.get_state()
    os.environ["NS2D_GRAPH"] = "0"
    try:
        with ns2d.Solver(cfg_of(p)) as s:
            s.set_state(w0)
            s.spectral_step(1e-3, 15)
            b = s.get_state()
    finally:
        del os.environ["NS2D_GRAPH"]
    assert np.array_equal(a, b)


def test_tuned_schedule_cache(monkeypatch, tmp_path):
    """The tuner's choice per size class is cached as JSON Lines keyed by
    device, shape tables, class size and member count: a first ensemble tunes
    and writes one record per class; a later ensemble with the same classes
    takes its shapes from the cache (here an edited record) without tuning,
    and still matches plain solves."""
    import json

    import ns2d
    cache = tmp_path / "sched.jsonl"
    monkeypatch.setenv("NS2D_ENS_CACHE", str(cache))
    monkeypatch.setenv("NS2D_ENS_TUNE", "1")
    params, w0s, dts = _members()
    cfgs = [cfg_of(p) for p in params]
    with ns2d.Ensemble(cfgs) as ens:
        for i, w in enumerate(w0s):
            ens[i].set_state(w)
        ens.step(dts, 2)
        tuned = [ens.variants(i) for i in range(len(ens))]
    recs = [json.loads(line) for line in cache.read_text().splitlines()]
    sizes = sorted({(p.Nx, p.Ny) for p in params})
    assert sorted((r["nx"], r["ny"]) for r in recs) == sizes
    for r in recs:
        i = next(j for j, p in enumerate(params) if (p.Nx, p.Ny) == (r["nx"], r["ny"]))
        assert (r["kpass"], r["rpass"]) == tuned[i]
        assert r["members"] == sum(1 for p in params if (p.Nx, p.Ny) == (r["nx"], r["ny"]))
    # a record naming other (valid) shapes: the next ensemble must use it
    edited = []
    for r in recs:
        r = dict(r, kpass=len(ns2d.kpass_variants(r["nx"])) - 1, rpass=len(ns2d.rpass_variants(r["ny"])) - 1)
        edited.append(r)
    with open(cache, "a") as f:
        for r in edited:
            f.write(json.dumps(r) + "\n")
    nsteps = 5
    with ns2d.Ensemble(cfgs) as ens:
        for i, w in enumerate(w0s):
            ens[i].set_state(w)
        ens.step(dts, nsteps)
        got = [ens[i].get_state() for i in range(len(ens))]
        used = [ens.variants(i) for i in range(len(ens))]
    for i, p in enumerate(params):
        r = next(r for r in edited if (r["nx"], r["ny"]) == (p.Nx, p.Ny))
        assert used[i] == (r["kpass"], r["rpass"]), i
        with ns2d.Solver(cfgs[i]) as s:
            s.set_state(w0s[i])
            s.spectral_step(dts[i], nsteps)
            plain = s.get_state()
        assert np.linalg.norm(got[i] - plain) / np.linalg.norm(plain) < 1e-13, i
