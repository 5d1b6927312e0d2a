"""MCMC densification restatements (oracle.mcmc_*: trainer.rs:1558-1650,
relocate.wgsl:39-125, mcmc.wgsl:25-58) pinned by known answers, invariants
and statistics, and the library's host sampler (gs_mcmc_sample_pairs, no
device work) against the restatement."""

import numpy as np

import oracle


def test_xorshift64_known_answer():
    # Marsaglia's (13, 7, 17) xorshift64 from state 1 (the reference's Trainer::rng step).
    x = oracle._xorshift64(1)
    assert x == 1082269761
    assert oracle._xorshift64(x) == 1152992998833853505


def _opacities(n=4000, seed=0):
    rng = np.random.default_rng(seed)
    op = rng.uniform(0.01, 1.0, n).astype(np.float32)
    op[rng.random(n) < 0.08] = np.float32(0.001)
    sc = np.exp(rng.normal(-4, 0.6, (n, 2))).astype(np.float32)
    return op, sc


def test_sampler_pairs_never_race_and_splits_take_leftover_slots():
    op, sc = _opacities()
    sc[17] = [1.0, 0.5]  # two oversized alive surfels
    sc[23] = [0.1, 2.0]
    op[17] = op[23] = 0.9
    pairs, rng = oracle.mcmc_sample_pairs(op, sc, 0.005, 8.0, 99)
    p = np.array(pairs)
    dead = set(np.nonzero(op < np.float32(0.005))[0].tolist())
    assert set(p[:, 0].tolist()) <= dead and not (set(p[:, 1].tolist()) & dead)
    assert len(set(p[:, 0].tolist())) == len(p) and len(set(p[:, 1].tolist())) == len(p)
    splits = p[p[:, 2] == 1]
    # The larger one first (2.0 > 1.0), each into a dead slot no relocation used.
    assert splits[0, 1] == 23 and splits[1, 1] == 17
    reloc = p[p[:, 2] == 0]
    unused = [d for d in sorted(dead) if d not in set(reloc[:, 0].tolist())]
    assert splits[0, 0] == unused[0] and splits[1, 0] == unused[1]
    assert rng != 99


def test_sampler_picks_alive_surfels_in_proportion_to_opacity():
    """The first pick of a batch lands on alive surfel i with probability
    opacity_i / sum (chi-square over 4000 batches, 5 alive surfels)."""
    op = np.array([0.001] * 3 + [0.1, 0.2, 0.3, 0.15, 0.25], np.float32)
    counts = np.zeros(8)
    rng = 7
    for _ in range(4000):
        pairs, rng = oracle.mcmc_sample_pairs(op, None, 0.005, 0.0, rng)
        counts[pairs[0][1]] += 1
    w = op[3:].astype(np.float64) / op[3:].sum()
    exp = 4000 * w
    chi2 = float(((counts[3:] - exp) ** 2 / exp).sum())
    assert counts[:3].sum() == 0 and chi2 < 20.0  # 4 dof: p ~ 5e-4


def test_host_sampler_matches_the_restatement():
    import splattar_amd as sa
    for seed, ratio in ((1, 8.0), (2, 0.0), (3, 3.0)):
        op, sc = _opacities(3000, seed)
        want, rw = oracle.mcmc_sample_pairs(op, sc, 0.005, ratio, 0x1234567 + seed)
        got, rg = sa.mcmc_sample_pairs(op, sc, 0.005, ratio, 0x1234567 + seed)
        assert rg == rw
        assert np.array_equal(got, np.array(want, np.uint32).reshape(-1, 4))
    e, r = sa.mcmc_sample_pairs(np.ones(10, np.float32), None, 0.005, 0.0, 5)
    assert e.shape == (0, 4) and r == 5  # no dead surfel: no draw


def test_split_opacity_composes_to_the_original():
    """alpha' = 1 - sqrt(1 - alpha) on both halves: two stacked halves
    transmit (1 - alpha')^2 = 1 - alpha, so the pair composites as the
    original surfel (the purpose of relocate.wgsl:50-54)."""
    raw = {"pos": np.zeros((4, 4), np.float32), "scales": np.full((4, 2), -3.0, np.float32),
           "quat": np.tile(np.array([0, 0, 0, 1], np.float32), (4, 1)), "opacity": np.zeros(4, np.float32),
           "sh": np.zeros((4, 48), np.float32)}
    m = {k: np.ones_like(v) for k, v in raw.items()}
    v = {k: np.ones_like(x) for k, x in raw.items()}
    act = np.array([0.001, 0.7, 0.002, 0.3], np.float32)
    oracle.mcmc_relocate_arrays([(0, 1, 0, 0), (2, 3, 1, 0)], act, raw, m, v)
    for d, a in ((0, 1), (2, 3)):
        an = 1.0 / (1.0 + np.exp(-np.float64(raw["opacity"][d])))
        assert abs((1 - (1 - an) ** 2) - act[a]) < 1e-6
        assert raw["opacity"][d] == raw["opacity"][a]
        for k in raw:
            assert np.all(m[k][d] == 0) and np.all(m[k][a] == 0) and np.all(v[k][a] == 0)
    assert np.allclose(raw["scales"][0], -3.16) and np.allclose(raw["scales"][2], -3.51)
    # The split halves sit +-0.6 of the major axis apart: 1.2 exp(-3) here.
    assert abs(np.linalg.norm(raw["pos"][2][:3] - raw["pos"][3][:3]) - 1.2 * np.exp(-3.0)) < 1e-6


def test_noise_statistics_and_gate():
    """Sum of four uniforms - 2: mean 0, variance 1/3; amplitude sigma x mean
    scale x exp(-k opacity); zero sigma moves nothing."""
    n = 200_000
    pos = np.zeros((n, 4), np.float32)
    sc = np.full((n, 2), 0.5, np.float32)
    op = np.zeros(n, np.float32)
    out = oracle.mcmc_noise(pos, sc, op, seed=123, sigma=1.0)
    d = out[:, :3].ravel().astype(np.float64) / 0.5
    assert abs(d.mean()) < 0.005 and abs(d.var() - 1.0 / 3.0) < 0.005 and np.abs(d).max() <= 2.0
    op2 = np.full(n, 0.4, np.float32)
    out2 = oracle.mcmc_noise(pos, sc, op2, seed=123, sigma=1.0)
    np.testing.assert_allclose(out2[:, :3], out[:, :3] * np.exp(-2.0), rtol=1e-5, atol=1e-12)
    assert np.array_equal(oracle.mcmc_noise(pos, sc, op, seed=123, sigma=0.0), pos)
