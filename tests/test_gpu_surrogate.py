"""GPU tests of the chemistry surrogate: the BF16 matrix-core kernels
against numpy on the same BF16 inputs, forward and gradients against the
FP64 oracle (oracle/mlp_oracle.py), Adam, conservation of the projected
substep, and what the trained surrogate gains over the BDF path alone.

Bars (written here): GEMM with FP32 accumulation of BF16 products, relative
1e-5 of max |C|; outputs stored as BF16 within 2^-8 relative; network
outputs and gradients with BF16 activations between layers within 3e-2
(relative Frobenius norm) of the FP64 oracle on the same BF16 weights and
inputs; accepted substeps keep every element's mass to 1e-12 relative;
the surrogate with its BDF fallback keeps max |T - T_BDF| <= 2 K on
held-out cells and runs more cells/s than the BDF path alone."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(mech):
    from phd_amd.ics import Box
    from phd_amd.solver import Options, Solver
    return Solver(Box(8, 8, 1, 3, 2, (1, 1, 1)), Options(), mech=mech)


def _bf16_bits(a):
    from oracle.mlp_oracle import bf16_round
    r = bf16_round(a).astype(np.float32)
    return (r.view(np.uint32) >> 16).astype(np.uint16).view(np.int16), r.astype(np.float64)


def _p(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_layouts_and_epilogues(mech, ta, tb):
    import torch
    from oracle.mlp_oracle import gelu, gelu_d
    s = _solver(mech)
    rng = np.random.default_rng(ta * 2 + tb)
    M, N, K = 100, 70, 45  # not multiples of the tile sizes
    A = rng.normal(size=(M, K)); B = rng.normal(size=(K, N))
    Ab, Af = _bf16_bits(A); Bb, Bf = _bf16_bits(B)
    Ast = Ab if ta == 0 else np.ascontiguousarray(Ab.T)
    Bst = np.ascontiguousarray(Bb.T) if tb == 0 else Bb
    lda = K if ta == 0 else M
    ldb = K if tb == 0 else N
    dA = torch.tensor(Ast, device="cuda"); dB = torch.tensor(Bst, device="cuda")
    bias = rng.normal(size=N).astype(np.float32)
    dbias = torch.tensor(bias, device="cuda")
    ref = Af @ Bf
    Cf = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    Cb = torch.zeros((M, N), dtype=torch.int16, device="cuda")
    lib = s.lib
    # plain product
    assert lib.izem_mlp_gemm(s.ctx, ta, tb, 0, M, N, K, _p(dA), lda, _p(dB), ldb, None, None, 0,
                             _p(Cf), N, None, 0) == 0
    s.sync()
    scale = np.abs(ref).max()
    assert np.max(np.abs(Cf.cpu().numpy() - ref)) <= 1e-5 * scale
    # bias + GELU (pre-activation fp32, activation bf16)
    assert lib.izem_mlp_gemm(s.ctx, ta, tb, 1, M, N, K, _p(dA), lda, _p(dB), ldb, _p(dbias),
                             None, 0, _p(Cf), N, _p(Cb), N) == 0
    s.sync()
    z = ref + bias
    assert np.max(np.abs(Cf.cpu().numpy() - z)) <= 1e-5 * np.abs(z).max()
    act = (Cb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    assert np.all(np.abs(act - gelu(z)) <= 2 ** -8 * np.abs(gelu(z)) + 1e-5 * np.abs(z).max())
    # GELU backward epilogue
    aux = torch.tensor(z.astype(np.float32), device="cuda")
    assert lib.izem_mlp_gemm(s.ctx, ta, tb, 3, M, N, K, _p(dA), lda, _p(dB), ldb, None,
                             _p(aux), N, None, 0, _p(Cb), N) == 0
    s.sync()
    g = (Cb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    gr = ref * gelu_d(z.astype(np.float32).astype(np.float64))
    assert np.all(np.abs(g - gr) <= 2 ** -8 * np.abs(gr) + 2e-5 * np.abs(gr).max())


def test_splitk_and_colsum(mech):
    import torch
    s = _solver(mech)
    rng = np.random.default_rng(5)
    M, N, K = 96, 64, 5000
    A = rng.normal(size=(K, M)); B = rng.normal(size=(K, N))  # both stored [K][*]
    Ab, Af = _bf16_bits(A); Bb, Bf = _bf16_bits(B)
    dA = torch.tensor(Ab, device="cuda"); dB = torch.tensor(Bb, device="cuda")
    out = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    work = torch.zeros(16 * M * N, dtype=torch.float32, device="cuda")
    assert s.lib.izem_mlp_gemm_splitk(s.ctx, 1, 1, M, N, K, _p(dA), M, _p(dB), N, 16, _p(work),
                                      _p(out)) == 0
    cs = torch.zeros(N, dtype=torch.float32, device="cuda")
    assert s.lib.izem_mlp_colsum(s.ctx, K, N, _p(dB), N, 16, _p(work), _p(cs)) == 0
    s.sync()
    ref = Af.T @ Bf
    assert np.max(np.abs(out.cpu().numpy() - ref)) <= 1e-5 * np.abs(ref).max() * 4
    assert np.max(np.abs(cs.cpu().numpy() - Bf.sum(axis=0))) <= 1e-4 * np.abs(Bf).sum(axis=0).max()


def test_forward_and_gradients_vs_fp64_oracle(mech):
    import torch
    from oracle import mlp_oracle as mo
    from phd_amd import surrogate as su
    s = _solver(mech)
    sur = su.Surrogate(s, seed=3)
    n = 512
    rng = np.random.default_rng(7)
    X = np.zeros((n, su.XLD)); X[:, :su.NF] = rng.normal(size=(n, su.NF))
    Xb, Xf = _bf16_bits(X)
    tgt = rng.normal(size=(n, su.NO)).astype(np.float32)
    dX = torch.tensor(Xb, device="cuda")
    dT = torch.tensor(tgt, device="cuda")
    pred = sur.forward(dX, keep=True).clone()
    dY = torch.zeros((n, 64), dtype=torch.int16, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    assert s.lib.izem_mlp_mse(s.ctx, n, su.NO, 64, _p(pred), 64, _p(dT), su.NO, _p(dY), 64,
                              _p(loss)) == 0
    grads = sur.backward(dY)
    s.sync()
    # the oracle on the same BF16 weights and inputs
    params = []
    for li, ((K, N), (k, nn)) in enumerate(zip(su.DIMS, su.TRUE_DIMS)):
        W = mo.bf16_round(sur.W[li].cpu().numpy().astype(np.float64))[:k, :nn]
        params.append((W, sur.b[li].cpu().numpy().astype(np.float64)[:nn]))
    out, _ = mo.forward(params, Xf[:, :su.NF])
    p = pred.cpu().numpy()[:, :su.NO]
    rel = np.linalg.norm(p - out) / np.linalg.norm(out)
    print("forward relative error", rel)
    assert rel <= 3e-2
    lref, gref = mo.loss_and_grads(params, Xf[:, :su.NF], tgt.astype(np.float64))
    assert abs(loss.item() / (n * su.NO) - lref) <= 3e-2 * lref
    for li, ((k, nn), (gW, gb)) in enumerate(zip(su.TRUE_DIMS, gref)):
        dW = grads[li][0].cpu().numpy()[:k, :nn]
        db = grads[li][1].cpu().numpy()[:nn]
        eW = np.linalg.norm(dW - gW) / np.linalg.norm(gW)
        eb = np.linalg.norm(db - gb) / np.linalg.norm(gb)
        print(f"layer {li}: dW rel {eW:.2e}, db rel {eb:.2e}")
        assert eW <= 3e-2 and eb <= 3e-2
        # padding stays exactly zero
        assert np.all(grads[li][0].cpu().numpy()[k:, :] == 0) and np.all(
            grads[li][0].cpu().numpy()[:, nn:] == 0)


def test_fused_inference_equals_layered_forward(mech):
    """The single-kernel network (activations in LDS) and the per-layer
    kernels accumulate in the same order: identical outputs."""
    import torch
    from phd_amd import surrogate as su
    s = _solver(mech)
    sur = su.Surrogate(s, seed=5)
    for li in range(4):  # nonzero biases
        sur.b[li].normal_(0.0, 0.1)
    n = 1000  # not a multiple of the 64-row tile
    rng = np.random.default_rng(8)
    X = np.zeros((n, su.XLD)); X[:, :su.NF] = rng.normal(size=(n, su.NF))
    Xb, _ = _bf16_bits(X)
    dX = torch.tensor(Xb, device="cuda")
    a = sur.forward(dX).clone()
    b = sur.infer(dX).clone()
    s.sync()
    assert torch.equal(a[:, :su.NO], b[:, :su.NO])


def test_adam_step_vs_oracle(mech):
    import torch
    from oracle import mlp_oracle as mo
    s = _solver(mech)
    rng = np.random.default_rng(9)
    rows, cols = 33, 70
    P = rng.normal(size=(rows, cols)).astype(np.float32)
    m = np.zeros_like(P); v = np.zeros_like(P)
    Pd, md, vd = (torch.tensor(a, device="cuda") for a in (P, m, v))
    Pt = torch.zeros((cols, rows), dtype=torch.int16, device="cuda")
    Pw = torch.zeros((rows, cols), dtype=torch.int16, device="cuda")
    Pr, mr, vr = P.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for t in range(1, 4):
        G = rng.normal(size=(rows, cols)).astype(np.float32)
        Gd = torch.tensor(G, device="cuda")
        assert s.lib.izem_mlp_adam(s.ctx, rows, cols, _p(Pd), _p(Gd), _p(md), _p(vd), 1e-2, 0.9,
                                   0.999, 1e-8, t, _p(Pt), rows, _p(Pw), cols) == 0
        mo.adam(Pr, G.astype(np.float64), mr, vr, t, 1e-2)
    s.sync()
    assert np.max(np.abs(Pd.cpu().numpy() - Pr)) <= 1e-5
    wb = (Pw.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    wt = (Pt.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(wb, wt.T)
    assert np.all(np.abs(wb - Pd.cpu().numpy()) <= 2 ** -8 * np.abs(Pd.cpu().numpy()))


def test_surrogate_trains_conserves_and_gains(mech, thd, adi):
    """Train on chemistry trajectories of reacting-HIT cells labelled by the
    BDF path, calibrate the indicator on held-out cells, then on test cells:
    element conservation of the accepted substeps, max |T - T_BDF| <= 2 K
    with the fallback, and the speed gain over the BDF path alone (both
    printed)."""
    import torch
    from phd_amd import surrogate as su
    s = _solver(mech)
    dts = [2e-8, 5e-8, 1e-7, 2e-7] * 3
    train = []
    for seed in (11, 12):
        train += su.trajectory(s, *su.hit_cells(adi, thd, 48, seed), dts)
    sur = su.Surrogate(s, seed=0)
    sur.set_normalisation(train)
    hist = sur.fit(train, epochs=20, batch=16384, lr=1e-3, verbose=True)
    assert hist[-1] < 0.5 * hist[0]
    cal = sur.calibrate(su.concat(su.trajectory(s, *su.hit_cells(adi, thd, 32, 13), [5e-8] * 6)))
    print("calibration", cal)
    L = su.concat(su.trajectory(s, *su.hit_cells(adi, thd, 48, 14), [5e-8] * 6))
    # element conservation of the accepted substeps
    T, Y = L.T.clone(), L.Y.clone()
    flag = sur.apply(T, L.rho, Y, L.dt)
    acc = (flag == 0).cpu().numpy()
    kn = np.asarray(thd.kncf, float); aw = np.asarray(thd.awt, float)
    E = kn * aw[:, None] / (aw @ kn)[None, :]
    e0 = L.Y.cpu().numpy() @ E.T
    e1 = Y.cpu().numpy() @ E.T
    assert acc.any()
    assert np.max(np.abs(e1[acc] - e0[acc])) <= 1e-12
    res = su.evaluate(sur, L)
    print("surrogate vs BDF:", res)
    assert res["max_abs_dT_K"] <= 2.0
    assert res["fallback_rate"] < 1.0
    assert res["surrogate_cells_per_s"] > res["vode_cells_per_s"]
