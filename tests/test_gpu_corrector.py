"""Learned corrector on the GPU: the BF16-MFMA product against the same
product on BF16-rounded operands (which also checks the MFMA operand lane
layout), the training kernels' loss and gradients against the CPU oracle,
the feature kernel against the oracle's neighbourhood, and what a short
training gains over plain bilinear interpolation on held-out directions."""

import math

import numpy as np
import pytest

import mlp_oracle as M
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


@pytest.mark.parametrize("M_,N_,K_", [(16, 16, 16), (37, 19, 45), (64, 64, 18), (64, 18, 4099)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_bf16(torch, M_, N_, K_, ta, tb):
    from tod_generation_mapbased_beam_amd import _lib
    from tod_generation_mapbased_beam_amd.device import stream_handle

    rng = np.random.default_rng(M_ * 7 + N_ + K_)
    A = rng.standard_normal((K_, M_) if ta else (M_, K_)).astype(np.float32)
    B = rng.standard_normal((N_, K_) if tb else (K_, N_)).astype(np.float32)
    C0 = rng.standard_normal((M_, N_)).astype(np.float32)
    dA, dB, dC = (torch.as_tensor(x, device="cuda") for x in (A, B, C0))
    _lib.check(_lib.lib().todg_corr_gemm_bf16(M_, N_, K_, _lib.ptr(dA), A.shape[1], ta, _lib.ptr(dB),
                                              B.shape[1], tb, _lib.ptr(dC), N_, 0.5, 2.0, stream_handle()))
    got = dC.cpu().numpy().astype(np.float64)
    opA = M.bf16(A.T if ta else A)
    opB = M.bf16(B.T if tb else B)
    ref = 0.5 * (opA @ opB) + 2.0 * C0
    scale = np.sqrt(K_) * 2.0
    assert np.max(np.abs(got - ref)) <= 1e-5 * scale


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.N_IN)).astype(np.float32)
    X[:, :M.GEO] = rng.uniform(-0.5, 0.5, (n, M.GEO))
    xh = X[:, M.GEO:].astype(np.float64)
    xh -= xh.mean(axis=1, keepdims=True)
    xh /= np.sqrt(np.mean(xh ** 2, axis=1, keepdims=True))
    X[:, M.GEO:] = xh
    w = rng.uniform(0.0, 1.0, (n, 4))
    w = (w / w.sum(axis=1, keepdims=True)).astype(np.float32)
    t = (rng.standard_normal(n) * 0.3).astype(np.float32)
    return X, w, t


def test_loss_and_gradient_match_oracle(torch):
    from tod_generation_mapbased_beam_amd.corrector import Corrector

    X, w, t = _batch(3000, 1)
    p = M.init_params(5)
    p = (p + np.random.default_rng(6).standard_normal(p.size).astype(np.float32) * 0.02)
    c = Corrector(params=p)
    loss, g, pred = c.loss_grad(*(torch.as_tensor(a, device="cuda") for a in (X, w, t)), want_pred=True)
    lo, go, po = M.loss_grad(p.astype(np.float64), X, w, t, bf16_ops=True)
    assert abs(float(loss[0]) - lo) <= 1e-4 * lo
    # bf16 rounding of an activation that sits on a rounding boundary in
    # float32 but not in float64 moves a prediction by ~1e-3: allow a few
    d = np.abs(pred.cpu().numpy() - po)
    assert np.quantile(d, 0.99) <= 2e-4 and d.max() <= 1e-2
    gg = g.cpu().numpy().astype(np.float64)
    from tod_generation_mapbased_beam_amd.corrector import N_H, N_IN, N_OUT

    # per parameter block of the flat layout W1 b1 W2 b2 W3 b3
    sizes = (N_H * N_IN, N_H, N_H * N_H, N_H, N_OUT * N_H, N_OUT)
    assert sum(sizes) == gg.size == go.size
    off = 0
    for n in sizes:
        a, b = gg[off:off + n], go[off:off + n]
        assert np.max(np.abs(a - b)) <= 2e-3 * np.max(np.abs(b)) + 1e-9, off
        off += n


def test_adam_step_matches_formula(torch):
    from tod_generation_mapbased_beam_amd.corrector import Corrector

    p = M.init_params(9)
    c = Corrector(params=p)
    g = np.random.default_rng(1).standard_normal(p.size).astype(np.float32)
    dg = torch.as_tensor(g, device="cuda")
    c.adam(dg, lr=1e-3)
    c.adam(dg, lr=1e-3)
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    w = p.astype(np.float64)
    for t in (1, 2):
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        w -= 1e-3 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    np.testing.assert_allclose(c.params.cpu().numpy(), w, atol=1e-6)


@pytest.mark.parametrize("nside", [8, 64])
def test_features_match_oracle_neighbourhood(torch, nside):
    from tod_generation_mapbased_beam_amd.corrector import features

    rng = np.random.default_rng(nside)
    n = 500
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * math.pi, n)
    field = rng.standard_normal(12 * nside * nside).astype(np.float32)
    F = features(nside, field, z, ph)
    pix = F.pix.cpu().numpy()
    ok = F.ok.cpu().numpy()
    X = F.X.cpu().numpy()
    wb = F.wbil.cpu().numpy()
    for i in range(n):
        ref = M.stencil16(O, nside, z[i], ph[i])
        assert (ref is not None) == bool(ok[i])
        if ref is None:
            continue
        p16, fa, fb, v, fm, fp = ref
        assert np.array_equal(pix[i], p16)
        m = field[p16].astype(np.float64)
        sd = np.sqrt(np.mean((m - m.mean()) ** 2))
        np.testing.assert_allclose(X[i, M.GEO:], (m - m.mean()) / sd, atol=1e-5)
        np.testing.assert_allclose(X[i, :M.GEO], [fa - 0.5, v - 0.5, fb - 0.5, fm - 0.5, fp - 0.5],
                                   atol=1e-6)
        np.testing.assert_allclose(wb[i], [(1 - v) * (1 - fa), (1 - v) * fa, v * (1 - fb), v * fb],
                                   atol=1e-6)


def test_training_beats_bilinear_on_held_out_sky(torch):
    """A short training on two nside-64 skies cuts the held-out squared
    residual against the bicubic judge well below plain bilinear's."""
    from tod_generation_mapbased_beam_amd import corrector as C

    X, W, T = C.make_pairs(64, seeds=[1, 2], n_per_map=100_000, lmax=191, dir_seed=1)
    Xh, Wh, Th = C.make_pairs(64, seeds=[99], n_per_map=50_000, lmax=191, dir_seed=2)
    c = C.Corrector(seed=0)
    c.train(X, W, T, epochs=4, batch=4096, lr=1e-3)
    bil = C.bilinear_loss(Xh, Wh, Th)
    got = c.eval_loss(Xh, Wh, Th)
    print(f"held-out mean sq residual: bilinear {bil:.4e}, corrector {got:.4e}")
    assert got < 0.5 * bil


def _mlp_infer(flat, X):
    """The fused kernel's MLP on rows X: bf16 operands with b1 folded into W1
    (so rounded to bf16 as well), FP32 b2, b3."""
    p = M.unpack(flat)
    Xa = np.concatenate([X, np.ones((X.shape[0], 1))], axis=1)
    W1a = np.concatenate([p["W1"], p["b1"][:, None]], axis=1)
    a1 = M.gelu(M.bf16(Xa) @ M.bf16(W1a).T)
    a2 = M.gelu(M.bf16(a1) @ M.bf16(p["W2"]).T + p["b2"])
    return M.bf16(a2) @ M.bf16(p["W3"]).T + p["b3"]


def _learned_oracle(vec, params_rot, nside, mp, bv, c_q, c_u, flat, bound=np.inf):
    """CPU restatement of todg_gather_accum_learned on the oracle's geometry:
    the corrector reads the first field outside the Q/U pair (else Q), and
    its weights serve every field, Q/U rotated by spin-2 per pixel."""
    B = params_rot[1].shape[0]
    S = vec.shape[0]
    C = mp.shape[0]
    d = O.rotated_dirs(vec, *params_rot, 0, B).astype(np.float64)
    tod = np.zeros((C, B))
    qu = c_q >= 0 and c_u >= 0
    others = [c for c in range(C) if not qu or c not in (c_q, c_u)]
    drv = others[0] if others else c_q
    for b in range(B):
        bx, by, bz = (float(x) for x in params_rot[3][b])
        z_pts = max(-1.0, min(1.0, bz))
        sth_pts = math.sqrt(max(0.0, 1.0 - bz * bz))
        phi_pts = math.atan2(by, bx) % (2 * math.pi)
        for s in range(S):
            z = max(-1.0, min(1.0, d[b, s, 2]))
            ph = math.atan2(d[b, s, 1], d[b, s, 0])
            ph = ph + 2 * math.pi if ph < 0 else ph
            st = M.stencil16(O, nside, z, ph)
            pix, w = O.ring_interp(nside, z, ph)[:2]
            if st is not None:
                p16, fa, fb, v, fm, fp = st
                m = mp[drv, p16].astype(np.float64)
                mu = m.mean()
                sd = math.sqrt(np.mean((m - mu) ** 2))
                xh = (m - mu) / sd if sd > 1e-30 else np.zeros(16)
                X = np.concatenate([[fa - 0.5, v - 0.5, fb - 0.5, fm - 0.5, fp - 0.5], xh])[None, :]
                o = _mlp_infer(flat, X)[0]
                dw = o - o.mean()
                if abs(float(np.dot(dw, xh))) <= bound:
                    w16 = dw.copy()
                    w16[[5, 6, 9, 10]] += [(1 - v) * (1 - fa), (1 - v) * fa, v * (1 - fb), v * fb]
                    pix, w = p16, w16
            vals = mp[:, pix].astype(np.float64).copy()
            if qu:
                for k, p in enumerate(pix):
                    zc, pc = O.pix2zphi(nside, int(p))
                    c2, s2 = O.spin2(zc, math.sqrt(max(0.0, 1 - zc * zc)), pc, z_pts, sth_pts, phi_pts)
                    q, u = vals[c_q, k], vals[c_u, k]
                    vals[c_q, k] = q * c2 + u * s2
                    vals[c_u, k] = -q * s2 + u * c2
            tod[:, b] += float(bv[s]) * (vals @ np.asarray(w, dtype=np.float64))
    return tod


def test_learned_tod_with_zero_correction_is_bilinear(torch):
    from cases import random_rot_params, rel_err
    from tod_generation_mapbased_beam_amd.corrector import gather_accum_learned

    rng = np.random.default_rng(21)
    nside, B, S = 32, 40, 300
    vec, params = random_rot_params(B, S, rng, pole_safe=False)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    zero = np.zeros(M.N_PARAMS, np.float32)
    for qu in ((1, 2), (-1, -1)):
        got = np.zeros((3, B))
        gather_accum_learned(vec, *params, nside, mp, bv, B, S, got, *qu, params=zero)
        ref = np.zeros((3, B))
        O.gather_accum_fused(vec, *params, nside, mp, bv, B, S, ref, *qu)
        assert rel_err(got, ref) <= 2e-6


@pytest.mark.parametrize("case", ["compact", "iqu", "iqu_bound", "i_only", "qu_only", "nside32"])
def test_learned_tod_matches_oracle_restatement(torch, case):
    """The fused footprint kernel against the CPU restatement (BF16 MLP): a
    compact beam on a ring scan (the certified fast path), and beams spread
    over the sphere (wide footprints, the exact path) for the channel
    layouts, a tight bound (many beam pixels fall back to bilinear) and an
    nside whose full-sphere footprint exceeds the LDS grid (direct
    accumulation)."""
    from cases import random_rot_params, rel_err
    from tod_generation_mapbased_beam_amd import inputs
    from tod_generation_mapbased_beam_amd import tod_rotations as R
    from tod_generation_mapbased_beam_amd.corrector import gather_accum_learned

    rng = np.random.default_rng(22)
    if case == "compact":
        nside, B = 64, 24
        vec, bv, _ = inputs.gaussian_beam(128, 120.0, 0.99)
        S = vec.shape[0]
        params = R.rotation_params_for_pointing(*inputs.ring_scan(B, 0.5))
    else:
        nside, B, S = (32, 4, 60) if case == "nside32" else (16, 6, 40)
        vec, params = random_rot_params(B, S, rng)
        bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
        bv /= bv.sum()
    C = {"i_only": 1, "qu_only": 2}.get(case, 3)
    qu = {"i_only": (-1, -1), "qu_only": (0, 1)}.get(case, (1, 2))
    bound = 0.05 if case == "iqu_bound" else np.inf
    mp = rng.standard_normal((C, 12 * nside * nside)).astype(np.float32)
    flat = M.init_params(3) * 4.0  # a visibly nonzero correction
    got = np.zeros((C, B))
    gather_accum_learned(vec, *params, nside, mp, bv, B, S, got, *qu, params=flat, bound=bound)
    ref = _learned_oracle(vec, params, nside, mp, bv, *qu, flat.astype(np.float64), bound)
    plain = np.zeros((C, B))
    O.gather_accum_fused(vec, *params, nside, mp, bv, B, S, plain, *qu)
    assert rel_err(got, ref) <= 5e-3
    if bound == np.inf:
        assert rel_err(got, plain) > 1e-2  # the correction is really applied


def test_learned_stats_and_bound(torch):
    """Counters add up to B*S; a zero bound with a nonzero correction falls
    back to plain bilinear everywhere (the result is the bilinear kernel's)."""
    from cases import random_rot_params, rel_err
    from tod_generation_mapbased_beam_amd.corrector import gather_accum_learned

    rng = np.random.default_rng(23)
    nside, B, S = 32, 30, 200
    vec, params = random_rot_params(B, S, rng)
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    bv = rng.uniform(0.1, 1.0, S).astype(np.float32)
    bv /= bv.sum()
    flat = M.init_params(5) * 4.0
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    got = np.zeros((3, B))
    gather_accum_learned(vec, *params, nside, mp, bv, B, S, got, 1, 2, params=flat, bound=0.0,
                         stats=stats)
    st = stats.cpu().numpy()
    assert st.sum() == B * S
    ref = np.zeros((3, B))
    O.gather_accum_fused(vec, *params, nside, mp, bv, B, S, ref, 1, 2)
    assert rel_err(got, ref) <= 2e-6


def test_learned_power_corrects_only_the_strongest_pixels(torch):
    """power < 1: the strongest beam pixels holding that power take the
    corrector and the rest plain bilinear, i.e. the TOD splits into a
    corrected part and a bilinear part over the two pixel sets."""
    from cases import rel_err
    from tod_generation_mapbased_beam_amd import inputs
    from tod_generation_mapbased_beam_amd import kernels as K
    from tod_generation_mapbased_beam_amd import tod_rotations as R
    from tod_generation_mapbased_beam_amd.corrector import gather_accum_learned

    rng = np.random.default_rng(31)
    nside, B = 64, 40
    vec, bv, _ = inputs.gaussian_beam(128, 120.0, 0.99)
    S = vec.shape[0]
    params = R.rotation_params_for_pointing(*inputs.ring_scan(B, 0.5))
    mp = rng.standard_normal((3, 12 * nside * nside)).astype(np.float32)
    flat = M.init_params(3) * 4.0
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    got = np.zeros((3, B))
    gather_accum_learned(vec, *params, nside, mp, bv, B, S, got, 1, 2, params=flat, power=0.9, stats=stats)
    order = np.argsort(-np.abs(bv), kind="stable")
    cum = np.cumsum(np.abs(bv[order]).astype(np.float64))
    k = int(np.searchsorted(cum, 0.9 * cum[-1])) + 1
    top, rest = order[:k], order[k:]
    ref = np.zeros((3, B))
    gather_accum_learned(vec[top], *params, nside, mp, bv[top], B, k, ref, 1, 2, params=flat)
    K._gather_accum_fused_jit(vec[rest], *params, nside, mp, bv[rest], B, S - k, ref, 1, 2)
    assert rel_err(got, ref) <= 1e-9
    st = stats.cpu().numpy()
    assert st.sum() == B * S and st[0] <= B * k and st[1] >= B * (S - k)
