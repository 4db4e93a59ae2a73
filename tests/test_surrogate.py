"""Learned near-field surrogate: oracle pins (CPU) and device parity (GPU).

CPU: the NumPy MLP oracle's gradient against central finite differences, its
BF16 rounding against hand-worked cases, the pair features' invariances and
the near-field sampler against a brute-force enumeration.
GPU (through the C ABI): forward and gradient of the BF16 MFMA kernels against
the oracle with the same BF16 rounding points (tolerance 2e-2 normwise per
parameter tensor, 1e-2 on predictions: what is left is FP32 accumulation
order, which flips single BF16 roundings); features and analytic labels
against the SA oracle (labels: 1e-6 of the largest label, the FP32 storage);
training reduces the loss; the assembled near-field blocks beat a
no-network predictor; out-of-box pairs fall back to the analytic routine bit
for bit.
"""
import math

import numpy as np
import pytest

from conftest import frob_rel


def _mlp():
    from oracle import mlp_oracle
    return mlp_oracle


# ----------------------------------------------------------------- CPU pins --

def test_bf16_rounding_cases():
    M = _mlp()
    one = 1.0
    cases = {
        one: one,
        one + 2.0 ** -8: one,                       # tie -> even (down)
        one + 3 * 2.0 ** -8: one + 2.0 ** -6,        # tie -> even (up)
        one + 2.0 ** -8 + 2.0 ** -20: one + 2.0 ** -7,  # above the tie
        -2.5: -2.5,
        3.0e-39: 3.0e-39,                             # subnormal, rounds within bf16's grid
    }
    for a, b in cases.items():
        r = float(M.bf16(np.float32(a)))
        if a == 3.0e-39:
            assert abs(r - a) <= 2.0 ** -133
        else:
            assert r == b, (a, r, b)


def test_mlp_oracle_gradient_matches_finite_differences():
    M = _mlp()
    rng = np.random.default_rng(3)
    p = M.random_params(1, scale=0.7).astype(np.float64)
    nm = M.random_norm(1)
    x = rng.normal(size=(6, M.N_IN)).astype(np.float32)
    y = rng.normal(size=(6, M.N_OUT)).astype(np.float32)

    # forward() casts to float32; evaluate the FP64 network on float64 params
    def forward64(pp):
        xs, _ = M.standardise(x, None, nm)
        a = np.zeros((x.shape[0], M.LAYER_K[0]))
        a[:, :M.N_IN] = xs
        o = 0
        for l, (K, N) in enumerate(zip(M.LAYER_K, M.LAYER_N)):
            W = pp[o:o + K * N].reshape(K, N)
            b = pp[o + K * N:o + K * N + N]
            o += K * N + N
            z = a @ W + b
            a = M.gelu(z) if l < 4 else z
        _, ys = M.standardise(x, y, nm)
        d = a[:, :M.N_OUT] - ys
        return float((d * d).sum() / d.size)

    loss, g = M.loss_grad(p.astype(np.float32), nm, x, y, emulate=False)
    assert abs(loss - forward64(p.astype(np.float32).astype(np.float64))) < 1e-12 * max(1.0, loss)
    p32 = p.astype(np.float32).astype(np.float64)
    idx = []
    for (W, b, wo, bo) in M.layers(np.zeros(M.PARAMS, np.float32)):
        K, N = W.shape
        kin = M.N_IN if wo == 0 else K
        nout = M.N_OUT if N == 64 else N
        for _ in range(4):
            idx.append(wo + rng.integers(kin) * N + rng.integers(nout))
        idx.append(bo + rng.integers(nout))
    for k in idx:
        e = 1e-5
        pp, pm = p32.copy(), p32.copy()
        pp[k] += e
        pm[k] -= e
        fd = (forward64(pp) - forward64(pm)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (k, fd, g[k])


def test_mlp_oracle_emulation_is_a_small_perturbation():
    M = _mlp()
    rng = np.random.default_rng(5)
    p = M.random_params(2)
    nm = M.random_norm(2)
    x = rng.normal(size=(64, M.N_IN)).astype(np.float32)
    a = M.predict(p, nm, x, emulate=True)
    b = M.predict(p, nm, x, emulate=False)
    r = frob_rel(a, b)
    assert 1e-5 < r < 5e-2, r  # rounding happens, and stays at BF16 size


def test_pair_features_are_rigid_and_scale_invariant():
    M = _mlp()
    rng = np.random.default_rng(7)
    Ts = rng.normal(size=(3, 3))
    Tf = rng.normal(size=(3, 3))
    x0 = M.pair_features(Ts, Tf, 5, 0.03, math.sqrt(3), 1.0, 0.4)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    t = rng.normal(size=3)
    x1 = M.pair_features(Ts @ Q.T + t, Tf @ Q.T + t, 5, 0.03, math.sqrt(3), 1.0, 0.4)
    np.testing.assert_allclose(x1, x0, rtol=0, atol=1e-12)
    x2 = M.pair_features(2.5 * Ts, 2.5 * Tf, 5, 2.5 * 0.03, math.sqrt(3), 1.0, 2.5 * 0.4)
    np.testing.assert_allclose(x2, x0, rtol=0, atol=1e-12)
    # field triangle in its own frame: centroid at 0, in the plane z = 0, first
    # edge along +e1, third vertex on the +e2 side
    F = x0[9:18].reshape(3, 3)
    assert np.allclose(F.sum(0), 0, atol=1e-12) and np.allclose(F[:, 2], 0, atol=1e-12)
    assert abs(F[0, 1] - F[1, 1]) < 1e-12 and F[1, 0] > F[0, 0] and F[2, 1] > F[0, 1]


def test_near_field_items_match_brute_force():
    from elastodynamicebem_amd.mesh import screen_uniform
    from elastodynamicebem_amd.surrogate import near_field_items
    m = screen_uniform(4)
    dt, cP, cS = 1 / 16, math.sqrt(3), 1.0
    L = 30
    got = set(zip(*[a.tolist() for a in near_field_items(m, dt, cP, cS, L)]))
    want = set()
    h = m.lMax
    for i in range(m.numberTriangles):
        for j in range(m.numberTriangles):
            if i == j:
                continue
            d = float(np.sqrt(((m.center[j] - m.center[i]) ** 2).sum()))
            if d > 3 * h:
                continue
            dmin = d - m.maxL[j] - m.maxL[i]
            dmax = d + m.maxL[j] + m.maxL[i]
            for lag in range(L):
                if (lag - 1) * cS * dt < dmax and (lag + 1) * cP * dt >= dmin:
                    want.add((i, j, lag))
    assert got == want and len(got) > 100


# --------------------------------------------------------------------- GPU --

def _near_samples(mesh, dt, n, seed=0):
    from elastodynamicebem_amd.surrogate import near_field_items
    i, j, g = near_field_items(mesh, dt, math.sqrt(3), 1.0)
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(i.size, min(n, i.size), replace=False))
    return i[sel], j[sel], g[sel]


@pytest.mark.gpu
def test_sur_predict_matches_oracle():
    from elastodynamicebem_amd.surrogate import Surrogate
    M = _mlp()
    rng = np.random.default_rng(11)
    p = M.random_params(4)
    nm = M.random_norm(4)
    x = rng.normal(size=(300, M.N_IN)).astype(np.float32)  # 300: padded to 320 rows
    s = Surrogate()
    try:
        s.set_params(p, nm)
        y = s.predict(x)
        ref = M.predict(p, nm, x, emulate=True)
        assert frob_rel(y, ref) < 1e-2, frob_rel(y, ref)
        # and against the unrounded network: BF16 size, not a layout error
        assert frob_rel(y, M.predict(p, nm, x, emulate=False)) < 5e-2
    finally:
        s.close()


@pytest.mark.gpu
def test_sur_gradient_matches_oracle():
    from elastodynamicebem_amd.surrogate import Surrogate, layer_offsets
    M = _mlp()
    rng = np.random.default_rng(12)
    p = M.random_params(5)
    nm = M.random_norm(5)
    x = rng.normal(size=(200, M.N_IN)).astype(np.float32)
    y = rng.normal(size=(200, M.N_OUT)).astype(np.float32)
    s = Surrogate()
    try:
        s.set_params(p, nm)
        g, loss = s.grad(x, y)
    finally:
        s.close()
    lref, gref = M.loss_grad(p, nm, x, y, emulate=True)
    assert abs(loss - lref) < 1e-3 * lref, (loss, lref)
    for l, (wo, bo, K, N) in enumerate(layer_offsets()):
        for a, b, what in ((wo, wo + K * N, "W"), (bo, bo + N, "b")):
            e = frob_rel(g[a:b], gref[a:b])
            assert e < 2e-2, (l, what, e)
    # padded rows/columns get exactly zero gradient
    wo, bo, K, N = layer_offsets()[0]
    assert not np.any(g[wo + M.N_IN * N:wo + K * N])
    wo, bo, K, N = layer_offsets()[-1]
    Wg = g[wo:wo + K * N].reshape(K, N)
    assert not np.any(Wg[:, M.N_OUT:]) and not np.any(g[bo + M.N_OUT:bo + N])


@pytest.mark.gpu
def test_sur_samples_match_sa_oracle(screen2, sphere2):
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.surrogate import Surrogate
    from oracle import Oracle
    M = _mlp()
    for mesh, Tfin, nT in ((screen2, 1.0, 16), (sphere2, 4.0, 128)):
        dt = Tfin / nT
        i, j, g = _near_samples(mesh, dt, 40, seed=1)
        pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=Tfin, nT=nT)
        s = Surrogate()
        try:
            with api.Solver(mesh, pb, 7) as sol:
                s.add_samples(sol, i, j, g)
            X, Y = s.get_samples()
        finally:
            s.close()
        o = Oracle(mesh.coordinates, mesh.triangles, Tfin=Tfin, nT=nT)
        h = mesh.lMax
        labels = []
        V = mesh.coordinates[mesh.triangles]
        for k in range(i.size):
            xf = M.pair_features(V[i[k]], V[j[k]], int(g[k]), dt, math.sqrt(3), 1.0, h)
            np.testing.assert_allclose(X[k], xf, rtol=1e-6, atol=1e-6)
            yo, Etot = M.pair_labels(o, int(i[k]), int(j[k]), int(g[k]), h)
            labels.append((Y[k], yo))
            Eb, ok = o.pair_block(int(g[k]), int(i[k]), int(j[k]))
            # P + S equals the analytic block (which zeroes entries below 1e-14)
            assert np.abs(Etot - Eb).max() <= 1e-13 + 1e-10 * np.abs(Eb).max()
        # labels: FP32 storage of FP64 values; labels that cancel to ~0 (a front
        # outside the pair at all three levels) are held to the mesh's label scale
        gmax = max(np.abs(yo).max() for _, yo in labels)
        for k, (y, yo) in enumerate(labels):
            err = np.abs(y - yo).max()
            assert err <= 1e-6 * np.abs(yo).max() + 1e-9 * gmax, (k, err, np.abs(yo).max(), gmax)


def _train_small(n_per_cfg=20000, epochs=6, only=None, seed=0):
    from elastodynamicebem_amd.surrogate import Surrogate, build_training_set, training_configs
    cfgs = [c for c in training_configs() if only is None or only(c.name)]
    s = Surrogate()
    ntr, nhe = build_training_set(s, n_per_cfg * len(cfgs), seed=seed, configs=cfgs, held_frac=0.1)
    n = s.num_samples()
    s.init(ntr, seed=seed)
    l0 = s.loss(ntr, n - ntr)
    losses = s.train(ntr, epochs, batch=1024, lr=2e-3, seed=seed)
    l1 = s.loss(ntr, n - ntr)
    return s, l0, losses, l1


@pytest.mark.gpu
def test_sur_training_and_assembly_gain(screen2):
    """Trained on screens, the network lowers the held-out loss and predicts
    the near-field blocks of the 8x8 screen far better than the network-free
    predictor (the training mean); the far-field blocks stay analytic and
    bit-identical."""
    from elastodynamicebem_amd import api
    s, l0, losses, l1 = _train_small(only=lambda nm: nm.startswith("screen"))
    try:
        assert losses[-1] < 0.5 * losses[0], losses
        assert l1 < 0.2 * l0, (l0, l1)
        pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=1.0, nT=16)
        with api.Solver(screen2, pb, 7) as sol:
            sol.assemble(api.PREC_FP64)
            ref = [sol.get_block(l) for l in range(sol.num_blocks())]
            s.attach(sol)
            sol.assemble(api.PREC_FP64 | api.PREC_SURROGATE)
            st = sol.surrogate_stats()
            got = [sol.get_block(l) for l in range(sol.num_blocks())]
            # network-free predictor: zero standardised output = training mean
            p, nm = s.params()
            s.set_params(np.zeros_like(p), nm)
            sol.assemble(api.PREC_FP64 | api.PREC_SURROGATE)
            mean_pred = [sol.get_block(l) for l in range(sol.num_blocks())]
            s.detach(sol)
        assert st["blocks"] > 0
        N = screen2.numberTriangles
        c = screen2.center
        d = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
        near = (d <= 3 * screen2.lMax) & ~np.eye(N, dtype=bool)
        mask = np.kron(near, np.ones((3, 3), bool))
        err_net = math.sqrt(sum(np.linalg.norm((g - r)[mask]) ** 2 for g, r in zip(got, ref)))
        err_mean = math.sqrt(sum(np.linalg.norm((m - r)[mask]) ** 2 for m, r in zip(mean_pred, ref)))
        nref = math.sqrt(sum(np.linalg.norm(r[mask]) ** 2 for r in ref))
        print(f"near-field rel error: network {err_net / nref:.3e}, mean predictor {err_mean / nref:.3e}")
        assert err_net < 0.5 * err_mean
        for g_, r in zip(got, ref):
            assert np.array_equal(g_[~mask], r[~mask])
    finally:
        s.close()


@pytest.mark.gpu
def test_sur_out_of_box_pairs_fall_back(screen2):
    """A network trained on spheres only never sees coplanar pairs (normal
    cosine 1, height 0): on the flat screen every pair falls back to the
    analytic routine and the result is bit-identical to the plain form."""
    from elastodynamicebem_amd import api
    s, *_ = _train_small(n_per_cfg=4000, epochs=1, only=lambda nm: nm.startswith("ico2"))
    try:
        pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=1.0, nT=16)
        with api.Solver(screen2, pb, 7) as sol:
            sol.assemble(api.PREC_FP64)
            ref = [sol.get_block(l) for l in range(sol.num_blocks())]
            s.attach(sol)
            sol.assemble(api.PREC_FP64 | api.PREC_SURROGATE)
            st = sol.surrogate_stats()
            got = [sol.get_block(l) for l in range(sol.num_blocks())]
            s.detach(sol)
        assert st["items"] == 0 and st["blocks"] == 0
        for g_, r in zip(got, ref):
            assert np.array_equal(g_, r)
    finally:
        s.close()


@pytest.mark.gpu
def test_sur_scatter_constant_network(sphere2):
    """Pins the prediction-to-block path on its own: a network whose output is
    a fixed label y0 (all weights zero, last bias y0, identity output
    normalisation, training box unbounded) must give, for every near-field
    pair (S, F) and every lag >= 1 of its window, the block
    R_F^T sym(y0_P + y0_S) R_F * A_S A_F / (4 pi rho h) -- R_F the field
    triangle's frame (first edge, normal x first edge, normal) -- to FP32
    precision; E_0 and every far-field block stay bit-identical to the
    analytic assembly, and the predicted lags form the pair's window without
    lag 0 (a lag off by one would move or drop one).  That is the raw
    prediction (EBEM_OPT_SURROGATE_PROJECT off); with the projection on, every
    predicted pair's blocks sum to zero over its lags and their first moment
    equals the plain form's (minus the static integral), as the exact blocks'
    do."""
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd import surrogate as SG
    rng = np.random.default_rng(3)
    y0 = rng.normal(size=SG.N_OUT).astype(np.float32)
    p = np.zeros(SG.PARAMS, np.float32)
    _, boff, _, _ = SG.layer_offsets()[-1]
    p[boff:boff + SG.N_OUT] = y0
    nm = np.zeros(SG.NORM, np.float32)
    nm[24:48] = 1.0            # x 1/std
    nm[48 + 18:48 + 36] = 1.0  # y std (y mean 0)
    nm[84:108] = -1e30         # box lo
    nm[108:132] = 1e30         # box hi
    m = sphere2
    pb = api.PbParam(rho=1.3, mu=1.0, lambda_=1.0, Tfin=4.0, nT=32)
    s = SG.Surrogate()
    try:
        s.set_params(p, nm)
        with api.Solver(m, pb, 7) as sol:
            sol.assemble(api.PREC_FP64)
            L = sol.num_blocks()
            ref = np.stack([sol.get_block(l) for l in range(L)])
            s.attach(sol)
            sol.set_option(api.OPT_SURROGATE_PROJECT, 0)
            sol.assemble(api.PREC_FP64 | api.PREC_SURROGATE)
            st = sol.surrogate_stats()
            got = np.stack([sol.get_block(l) for l in range(L)])
            sol.set_option(api.OPT_SURROGATE_PROJECT, 1)
            sol.assemble(api.PREC_FP64 | api.PREC_SURROGATE)
            proj = np.stack([sol.get_block(l) for l in range(L)])
            s.detach(sol)
    finally:
        s.close()
    lagv = np.arange(L)[:, None, None]
    n3 = proj.shape[1]
    blk0 = lambda A: A.reshape(A.shape[0], n3 // 3, 3, n3 // 3, 3)
    S0p = blk0(proj).sum(0)
    S1p, S1r = blk0(lagv * proj).sum(0), blk0(lagv * ref).sum(0)
    Sr = np.abs(blk0(ref)).max(0)  # scale of each pair's blocks
    N = m.numberTriangles
    V = m.coordinates[m.triangles]
    c = V.mean(1)
    d = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    h = m.lMax
    near = (d <= 3 * h * (1 - 1e-12)) & ~np.eye(N, dtype=bool)
    tol = (1e-12 * np.abs(S1r).max() + 1e-9 * Sr).transpose(0, 2, 1, 3)[near]
    assert np.all(np.abs(S0p.transpose(0, 2, 1, 3)[near]) <= tol)
    assert np.all(np.abs((S1p - S1r).transpose(0, 2, 1, 3)[near]) <= tol)
    assert st["blocks"] > 0
    # E_0 analytic (level 1 evaluated level by level, the plain form may have
    # taken it from a kP range: same value up to that split's rounding)
    assert np.abs(got[0] - ref[0]).max() <= 1e-13 * np.abs(ref[0]).max()
    blk = lambda A, i, j: A[:, 3 * i:3 * i + 3, 3 * j:3 * j + 3]
    M = (y0[:9] + y0[9:]).reshape(3, 3).astype(np.float64)
    Ms = 0.5 * (M + M.T)
    area = m.area
    checked = 0
    for i in range(N):
        for j in range(N):
            G, R0 = blk(got, i, j), blk(ref, i, j)
            if not near[i, j]:
                if d[i, j] > 3 * h * (1 + 1e-12):
                    assert np.array_equal(G, R0), (i, j)
                continue
            e1 = V[j, 1] - V[j, 0]
            e1 /= np.linalg.norm(e1)
            n = m.normal[j]
            R = np.array([e1, np.cross(n, e1), n])
            want = R.T @ Ms @ R * area[i] * area[j] / (4 * np.pi * pb.rho * h)
            lags = [l for l in range(1, L) if np.any(G[l] != 0)]
            ref_lags = [l for l in range(1, L) if np.any(R0[l] != 0)]
            if not lags:
                assert not ref_lags, (i, j)
                continue
            assert lags == list(range(lags[0], lags[-1] + 1)), (i, j, lags)
            assert set(ref_lags) <= set(lags), (i, j, ref_lags, lags)
            for l in lags:
                assert np.abs(G[l] - want).max() <= 1e-6 * np.abs(want).max(), (i, j, l)
            checked += 1
    assert checked > 100
