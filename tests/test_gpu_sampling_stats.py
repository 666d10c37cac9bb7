"""Frequencies of the sampler variants that no seeded sample list of the
reference covers, against probabilities from the pinned oracles.

The lockstep sampler follows the reference's draw sequence
(passive/sampling.py); the reference's own seeded sample lists pin the
ideal and lossy variants draw for draw (test_gpu_sampling.py).  For uniform
partial distinguishability and for postselection (lossless and lossy) there
is no such list, so here the empirical distribution of 10^5 shots on a
3-photon, 4-mode case is held to the exact output distribution with a
chi-square test: uniform overlap against the tensor-permanent oracle
(oracle.pd_detection_probability), postselection against |perm|^2 from the
double-double oracle and the lossy polynomial oracle
(oracle.lossy_pd_probabilities_dd), conditioned on the postselected modes.
Postselection with uniform overlap runs shot by shot (its restarts depend
on the draws, passive/sampling.py:422-529), so it gets 2 x 10^4 shots.
Seeds are fixed, so each test is deterministic; each also checks that the
same counts reject a plausible wrong distribution (the test has power).
"""

import itertools
import math

import numpy as np
import pytest
from scipy.stats import chi2

import piquasso_amd as pa
from piquasso_amd import workloads as W

pytestmark = pytest.mark.gpu

SHOTS = 100_000
P_MIN = 1e-3     # the exact distribution passes at this level
P_WRONG = 1e-8   # a wrong one fails below this


def _basis(d, n):
    return [s for s in itertools.product(range(n + 1), repeat=d) if sum(s) == n]


def _counts(samples, outcomes):
    index = {o: k for k, o in enumerate(outcomes)}
    c = np.zeros(len(outcomes))
    for s in samples:
        c[index[tuple(s)]] += 1
    return c


def _pvalue(counts, probs):
    """Chi-square p-value of counts against probabilities (summing to 1);
    outcomes expected fewer than 5 times are pooled into one bin."""
    shots = counts.sum()
    exp = probs * shots
    big = exp >= 5
    obs = list(counts[big])
    ex = list(exp[big])
    if (~big).any() and exp[~big].sum() > 0:
        obs.append(counts[~big].sum())
        ex.append(exp[~big].sum())
    obs, ex = np.array(obs), np.array(ex)
    stat = float(np.sum((obs - ex) ** 2 / ex))
    return float(chi2.sf(stat, len(ex) - 1)), stat, len(ex) - 1


def _unitary():
    return W.haar_unitary(4, np.random.default_rng(2024))


def test_uniform_overlap_frequencies(gpu, oracle):
    u = _unitary()
    inp = (1, 1, 1, 0)
    x = 0.6
    outs = _basis(4, 3)
    probs = np.array([oracle.pd_detection_probability(o, u, inp, x) for o in outs])
    assert abs(probs.sum() - 1.0) < 1e-12
    smp = pa.generate_samples(inp, SHOTS, u, uniform_particle_overlap=x, seed_sequence=11)
    counts = _counts(smp, outs)
    p, stat, df = _pvalue(counts, probs)
    print(f"uniform overlap x={x}: chi2 {stat:.1f} on {df} dof, p {p:.3g}", flush=True)
    assert p > P_MIN
    # the same counts against indistinguishable or distinguishable photons
    for wrong in (1.0, 0.0):
        pw = np.array([oracle.pd_detection_probability(o, u, inp, wrong) for o in outs])
        assert _pvalue(counts, pw)[0] < P_WRONG, wrong


def _postselected_probs(full_probs, outs_full, ps_modes, ps_photons):
    """Probabilities of the free modes' outcomes given the postselection,
    keyed by the outcome with the postselected modes removed."""
    keep = [k for k, o in enumerate(outs_full)
            if all(o[m] == c for m, c in zip(ps_modes, ps_photons))]
    free = [tuple(v for i, v in enumerate(outs_full[k]) if i not in ps_modes) for k in keep]
    p = np.array([full_probs[k] for k in keep])
    return free, p / p.sum()


def test_postselection_frequencies(gpu, oracle):
    u = _unitary()
    inp = (1, 1, 1, 0)
    ps_modes, ps_photons = (3,), (1,)
    outs_full = _basis(4, 3)
    full = np.array([abs(oracle.perm_dd(u, o, inp)) ** 2 / np.prod([math.factorial(k) for k in o])
                     for o in outs_full])
    assert abs(full.sum() - 1.0) < 1e-12
    outs, probs = _postselected_probs(full, outs_full, ps_modes, ps_photons)
    smp = pa.generate_samples(inp, SHOTS, u, postselect_data=(ps_modes, ps_photons, 1000),
                              seed_sequence=12)
    counts = _counts(smp, outs)
    p, stat, df = _pvalue(counts, probs)
    print(f"postselection: chi2 {stat:.1f} on {df} dof, p {p:.3g}", flush=True)
    assert p > P_MIN
    # against distinguishable photons (no interference), conditioned the same way
    fw = np.array([oracle.pd_detection_probability(o, u, inp, 0.0) for o in outs_full])
    assert _pvalue(counts, _postselected_probs(fw, outs_full, ps_modes, ps_photons)[1])[0] < P_WRONG


def test_lossy_postselection_frequencies(gpu, oracle):
    """generate_lossy_samples with postselection samples the 2d-mode unitary
    dilation B of T and, as the reference does (passive/sampling.py:110-146),
    trims each sample to its first d entries after the postselected modes
    are removed: the free modes of T, then as many auxiliary (loss) modes as
    there are postselected modes.  Its frequencies are held to |perm(B)|^2
    from the double-double oracle, marginalised and conditioned that way;
    the free modes' marginal is also held to T's lossy probabilities from
    the polynomial oracle (the dilation is the right one)."""
    from piquasso_amd.sampling import expanded_lossy_matrix

    u = _unitary()
    t = u @ np.diag([0.9, 0.8, 0.7, 0.95]) @ W.haar_unitary(4, np.random.default_rng(7))
    d = 4
    inp = (1, 1, 1, 0)
    ps_modes, ps_photons = (0,), (1,)
    big = expanded_lossy_matrix(t)
    inp_big = inp + (0,) * d
    outs_big = _basis(2 * d, 3)
    full = np.array([abs(oracle.perm_dd(big, o, inp_big)) ** 2 / np.prod([math.factorial(k) for k in o])
                     for o in outs_big])
    assert abs(full.sum() - 1.0) < 1e-12
    # the trimmed sample: postselected modes removed, then the first d entries
    probs = {}
    for o, q in zip(outs_big, full):
        if all(o[m] == c for m, c in zip(ps_modes, ps_photons)):
            key = tuple(v for i, v in enumerate(o) if i not in ps_modes)[:d]
            probs[key] = probs.get(key, 0.0) + q
    outs = sorted(probs)
    p_exact = np.array([probs[o] for o in outs])
    p_exact /= p_exact.sum()
    smp = pa.generate_lossy_samples(inp, SHOTS, t, postselect_data=(ps_modes, ps_photons, 1000),
                                    seed_sequence=13)
    counts = _counts(smp, outs)
    p, stat, df = _pvalue(counts, p_exact)
    print(f"lossy postselection: chi2 {stat:.1f} on {df} dof, p {p:.3g}", flush=True)
    assert p > P_MIN
    # the free modes' marginal against T's lossy probabilities (polynomial oracle)
    outs_t = [o for k in range(4) for o in _basis(d, k)]
    pt, _ = oracle.lossy_pd_probabilities_dd(np.array(outs_t), t, inp, 1.0, conjugate_loss=True)
    assert abs(pt.sum() - 1.0) < 1e-12
    free_t, pf = _postselected_probs(pt, outs_t, ps_modes, ps_photons)
    n_free = d - len(ps_modes)
    marg = {}
    for o, q in zip(outs, p_exact):
        marg[o[:n_free]] = marg.get(o[:n_free], 0.0) + q
    for o, q in zip(free_t, pf):
        assert abs(marg.get(o, 0.0) - q) < 1e-12, o
    # against the lossless unitary part of T (losses ignored)
    wrong = {}
    for o in outs_t:
        if sum(o) == 3 and o[0] == 1:
            key = tuple(v for i, v in enumerate(o) if i not in ps_modes) + (0,) * len(ps_modes)
            wrong[key] = abs(oracle.perm_dd(u, o, inp)) ** 2 / np.prod([math.factorial(k) for k in o])
    pw = np.array([wrong.get(o, 0.0) for o in outs])
    assert _pvalue(counts, pw / pw.sum())[0] < P_WRONG


def test_postselection_with_overlap_frequencies(gpu, oracle):
    """Postselection with uniformly partially distinguishable photons: the
    free modes' outcomes given the postselected ones, against the tensor-
    permanent oracle's probabilities conditioned the same way."""
    u = _unitary()
    inp = (1, 1, 1, 0)
    x = 0.6
    ps_modes, ps_photons = (3,), (1,)
    outs_full = _basis(4, 3)
    full = np.array([oracle.pd_detection_probability(o, u, inp, x) for o in outs_full])
    outs, probs = _postselected_probs(full, outs_full, ps_modes, ps_photons)
    shots = 20_000
    smp = pa.generate_samples(inp, shots, u, postselect_data=(ps_modes, ps_photons, 1000),
                              uniform_particle_overlap=x, seed_sequence=14)
    counts = _counts(smp, outs)
    p, stat, df = _pvalue(counts, probs)
    print(f"postselection, overlap x={x}: chi2 {stat:.1f} on {df} dof, p {p:.3g}", flush=True)
    assert p > P_MIN
    for wrong in (1.0, 0.0):
        fw = np.array([oracle.pd_detection_probability(o, u, inp, wrong) for o in outs_full])
        pw = _postselected_probs(fw, outs_full, ps_modes, ps_photons)[1]
        assert _pvalue(counts, pw)[0] < P_WRONG, wrong
