"""Adjoint kernels (iclblasSgemmBackward, iclblasStrsmBackward) through the
torch.autograd.Function harness (clgpu_amd.autograd), checked against
  * FP64 central finite differences of the loss evaluated with the FP64
    oracle (dgemm + dtrsm), on the 64 x 64 A / 64 x 16 B instance of the
    differentiable chain X = op(A)^-1 (alpha G H), loss = sum(W * X);
  * gradients assembled in FP64 from the CPU oracle's formulas for every
    transpose / side / triangle variant;
and the full-size chain (A 2048^2, G 2048^2, H 2048 x 512) forward + backward
on the GPU with sampled FP64 checks."""
import numpy as np
import pytest

import oracle
from clgpu_amd import datagen

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).T)).cuda().requires_grad_(True)


def host(t):
    return np.asfortranarray(t.detach().cpu().numpy().T.astype(np.float64))


def _chain_inputs(order, nrhs, seed=0):
    A = datagen.triangular(order, lower=True, seed=seed, poison=False)
    G = datagen.matrix(order, order, name="G", seed=seed)
    H = datagen.matrix(order, nrhs, name="H", seed=seed)
    W = np.asfortranarray(datagen.normal(order * nrhs, seed=seed, stream=5).reshape(nrhs, order).T)
    return A, G, H, W


def _loss_fp64(A, G, H, W, alpha):
    """loss = sum(W * X), X = A^-1 (alpha G H), all in float64 (numpy + FP64 substitution)."""
    import scipy.linalg as sla
    Y = alpha * (G @ H)
    X = sla.solve_triangular(np.tril(A), Y, lower=True)
    return float(np.sum(W * X))


def test_chain_fd_check_64x64():
    import torch
    from clgpu_amd import autograd as ag
    order, nrhs, alpha = 64, 16, 0.5
    A, G, H, W = _chain_inputs(order, nrhs)
    tA, tG, tH = dev(A), dev(G), dev(H)
    Y = ag.sgemm(tG, tH, alpha)
    X = ag.strsm(tA, Y, 1.0, 0, 1, 0, 0)
    loss = (X * torch.from_numpy(np.ascontiguousarray(W.T.astype(np.float32))).cuda()).sum()
    loss.backward()
    torch.cuda.synchronize()
    gA, gG, gH = host(tA.grad), host(tG.grad), host(tH.grad)

    A64, G64, H64 = A.astype(np.float64), G.astype(np.float64), H.astype(np.float64)
    eps = 1e-6

    def fd(which, shape, mask=None):
        g = np.zeros(shape)
        for j in range(shape[1]):
            for i in range(shape[0]):
                if mask is not None and not mask[i, j]:
                    continue
                args = [A64.copy(), G64.copy(), H64.copy()]
                args[which][i, j] += eps
                lp = _loss_fp64(*args, W, alpha)
                args[which][i, j] -= 2 * eps
                lm = _loss_fp64(*args, W, alpha)
                g[i, j] = (lp - lm) / (2 * eps)
        return g

    lower = np.tril(np.ones((order, order), bool))
    fA = fd(0, A.shape, lower)
    fG = fd(1, G.shape)
    fH = fd(2, H.shape)
    for got, ref in ((gA, fA), (gG, fG), (gH, fH)):
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel <= 1e-4, rel
    assert np.all(gA[~lower] == 0.0)  # gradient confined to the stored triangle


def _tri_clean(order, uplo, diag, seed):
    A = datagen.triangular(order, lower=bool(uplo), seed=seed, poison=False)
    return A


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [0, 1])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("diag", [0, 1])
def test_strsm_backward_vs_fp64_formulas(side, uplo, trans, diag):
    import torch
    from clgpu_amd import autograd as ag
    m, n, alpha = 70, 33, -0.8
    order = m if side == 0 else n
    A = _tri_clean(order, uplo, diag, seed=51)
    B = datagen.matrix(m, n, name="B", seed=51)
    W = datagen.matrix(m, n, name="W", seed=51)
    tA, tB = dev(A), dev(B)
    X = ag.strsm(tA, tB, alpha, side, uplo, trans, diag)
    X.backward(torch.from_numpy(np.ascontiguousarray(W.T)).cuda())
    torch.cuda.synchronize()
    # FP64 assembly from the oracle's substitution (the reference's Strsm_naive loops in double):
    #   X = alpha op(A)^-1 B,  G = op(A)^-T W  (the solve with the opposite trans),  dM = -G X^T / -X^T G
    X64 = oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    Gm = oracle.dtrsm(side, uplo, 1 - trans, diag, m, n, 1.0, A, W)
    dM = -Gm @ X64.T if side == 0 else -X64.T @ Gm
    dA = dM if trans == 0 else dM.T
    keep = np.tril(np.ones_like(dA, bool)) if uplo == 1 else np.triu(np.ones_like(dA, bool))
    if diag == 1:
        np.fill_diagonal(keep, False)
    dA = np.where(keep, dA, 0.0)
    dB = alpha * Gm
    gA, gB = host(tA.grad), host(tB.grad)
    assert np.linalg.norm(host(X) - X64) / np.linalg.norm(X64) < 1e-5
    assert np.linalg.norm(gB - dB) / np.linalg.norm(dB) < 1e-5
    assert np.linalg.norm(gA - dA) / np.linalg.norm(dA) < 1e-5
    assert np.all(gA[~keep] == 0.0)


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
def test_sgemm_backward_vs_fp64_formulas(ta, tb):
    import torch
    from clgpu_amd import autograd as ag
    m, n, k, alpha = 45, 70, 90, 1.5
    A = datagen.matrix(*((m, k) if ta == 0 else (k, m)), name="A", seed=61)
    B = datagen.matrix(*((k, n) if tb == 0 else (n, k)), name="B", seed=61)
    dC = datagen.matrix(m, n, name="C", seed=61)
    tA, tB = dev(A), dev(B)
    C = ag.sgemm(tA, tB, alpha, ta, tb)
    C.backward(torch.from_numpy(np.ascontiguousarray(dC.T)).cuda())
    torch.cuda.synchronize()
    C64, _ = oracle.dgemm(ta, tb, m, n, k, alpha, A, B)
    assert np.allclose(host(C), C64, rtol=1e-5, atol=1e-4)
    dopA, _ = oracle.dgemm(0, 1 - tb, m, k, n, alpha, dC, B)  # alpha * dC * op(B)^T
    dopB, _ = oracle.dgemm(1 - ta, 0, k, n, m, alpha, A, dC)  # alpha * op(A)^T * dC
    dA = dopA if ta == 0 else dopA.T
    dB = dopB if tb == 0 else dopB.T
    assert np.linalg.norm(host(tA.grad) - dA) / np.linalg.norm(dA) < 1e-5
    assert np.linalg.norm(host(tB.grad) - dB) / np.linalg.norm(dB) < 1e-5


def test_chain_config5_full_size():
    """X = A^-1 (alpha G H) with A 2048^2, G 2048^2, H 2048 x 512; forward + backward on
    the GPU; dG, dH and dA checked in FP64 on sampled entries."""
    import torch
    from clgpu_amd import autograd as ag
    order, nrhs, alpha = 2048, 512, 0.5
    A, G, H, W = _chain_inputs(order, nrhs, seed=0)
    tA, tG, tH = dev(A), dev(G), dev(H)
    Y = ag.sgemm(tG, tH, alpha)
    X = ag.strsm(tA, Y, 1.0, 0, 1, 0, 0)
    tW = torch.from_numpy(np.ascontiguousarray(W.T.astype(np.float32))).cuda()
    (X * tW).sum().backward()
    torch.cuda.synchronize()
    # FP64 reference on a sample of columns of the right-hand side
    import scipy.linalg as sla
    A64, G64, H64, W64 = (x.astype(np.float64) for x in (A, G, H, W))
    L = np.tril(A64)
    Y64 = alpha * G64 @ H64
    X64 = sla.solve_triangular(L, Y64, lower=True)
    dY64 = sla.solve_triangular(L.T, W64, lower=False)
    rng = np.random.default_rng(1)
    r = rng.integers(0, order, 200)
    c = rng.integers(0, order, 200)
    keep = r >= c
    dA_s = -(dY64[r] * X64[c]).sum(axis=1) * keep
    gA = host(tA.grad)[r, c]
    assert np.linalg.norm(gA - dA_s) / np.linalg.norm(dA_s) < 1e-4
    dH64 = alpha * G64.T @ dY64
    assert np.linalg.norm(host(tH.grad) - dH64) / np.linalg.norm(dH64) < 1e-4
    cols = rng.integers(0, order, 16)
    dG_s = alpha * dY64 @ H64[cols].T
    assert np.linalg.norm(host(tG.grad)[:, cols] - dG_s) / np.linalg.norm(dG_s) < 1e-4
    assert np.linalg.norm(host(X) - X64) / np.linalg.norm(X64) < 1e-4
