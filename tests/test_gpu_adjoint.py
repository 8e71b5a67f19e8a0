"""Adjoint kernel (spsim_hip_sf_adjoint) and the torch.autograd front end:
gradients of a Poisson NLL w.r.t. positions and per-atom B factors against the
FP64 oracle gradient (itself pinned by finite differences) and against a dense
FP64 PyTorch autograd reference."""
import numpy as np
import pytest

import oracle
from spsim_amd import device as dv
from spsim_amd import torch_ops as to
from spsim_amd import workloads as wl
from tests.conftest import oracle_threads, rel_l2

pytestmark = pytest.mark.gpu


def small_problem(natoms=120, n=24, seed=0, with_b=True):
    rng = np.random.default_rng(seed)
    Z, pos, B = wl.random_molecule(natoms, 2.5e-9, rng, atom_b_range=(10.0, 40.0) if with_b else None)
    det = wl.Detector(n, 1e-3, 0.1, 0.15e-9)
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    return Z, pos, B, q


def test_adjoint_matches_oracle_gradient(gpu):
    import torch

    Z, pos, B, q = small_problem()
    Z[5] = 0  # skipped atom: zero gradient
    rng = np.random.default_rng(1)
    g = rng.normal(size=q.shape[0])
    gp_ref, gb_ref = oracle.sf_grad_f64(q, Z, pos, g, atom_b=B, B=2.0, threads=oracle_threads())
    layout = to.SpeciesLayout(Z, gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    P = torch.from_numpy(pos).to(gpu).requires_grad_(True)
    Bt = torch.from_numpy(B).to(gpu).requires_grad_(True)
    I = to.intensity(P, hkl, layout, ff, atom_b=Bt, b_global=2.0)
    (I * torch.from_numpy(g.astype(np.float32)).to(gpu)).sum().backward()
    gp = P.grad.cpu().numpy()
    gb = Bt.grad.cpu().numpy()
    assert rel_l2(gp, gp_ref) < 2e-5
    assert rel_l2(gb, gb_ref) < 2e-5
    assert np.all(gp[5] == 0) and gb[5] == 0


def test_nufft_forward_with_adjoint_backward(gpu):
    """method="nufft": the intensity comes from the non-uniform FFT (per-atom B by
    Gaussian spreading) and the backward is the same adjoint fed with that F.
    Intensities and gradients match the FP64 oracle; an unknown method is refused."""
    import torch

    Z, pos, B, _ = small_problem(natoms=400, seed=3)
    q = oracle.hkl_detector(40, 40, 40 * 1e-3 / 24, 40 * 1e-3 / 24, 0.1, 0, 0, 0, 0.15e-9)
    rng = np.random.default_rng(4)
    g = rng.normal(size=q.shape[0])
    layout = to.SpeciesLayout(Z, gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    P = torch.from_numpy(pos).to(gpu).requires_grad_(True)
    Bt = torch.from_numpy(B).to(gpu).requires_grad_(True)
    I = to.intensity(P, hkl, layout, ff, atom_b=Bt, b_global=2.0, method="nufft")
    (I * torch.from_numpy(g.astype(np.float32)).to(gpu)).sum().backward()
    _, I_ref = oracle.sf_f64(q, Z, pos, atom_b=B, B=2.0, threads=oracle_threads())
    assert rel_l2(I.detach().cpu().numpy(), I_ref) < 1e-5
    gp_ref, gb_ref = oracle.sf_grad_f64(q, Z, pos, g, atom_b=B, B=2.0, threads=oracle_threads())
    assert rel_l2(P.grad.cpu().numpy(), gp_ref) < 1e-4
    assert rel_l2(Bt.grad.cpu().numpy(), gb_ref) < 1e-4
    with pytest.raises(ValueError, match="method"):
        to.intensity(P, hkl, layout, ff, method="fft")


def dense_torch_intensity(q, Z, pos, B, ff, torch):
    """FP64 dense reference on the CPU: I_p = |sum_j f_Zj(q_p) e^{-B_j s_p^2} e^{-2 pi i q_p.r_j}|^2."""
    qd = torch.from_numpy(q.astype(np.float64))
    d = torch.linalg.norm(qd, dim=1)
    s2 = (d * 1e-10) ** 2 / 4
    f = torch.zeros((q.shape[0], len(Z)), dtype=torch.float64)
    for z in np.unique(Z):
        c = torch.from_numpy(ff[z].astype(np.float64))
        fz = sum(c[i] * torch.exp(-c[4 + i] * s2) for i in range(4)) + c[8]
        f[:, torch.from_numpy(np.nonzero(Z == z)[0])] = fz[:, None]
    a = f * torch.exp(-B[None, :] * s2[:, None])
    ph = 2 * np.pi * (qd @ pos.T)
    re = (a * torch.cos(ph)).sum(1)
    im = -(a * torch.sin(ph)).sum(1)
    return re * re + im * im


def test_config4_gradient_vs_dense_fp64_torch(gpu):
    """Differentiation configuration at N = 500: Poisson NLL of a pattern
    simulated at perturbed coordinates against counts drawn at the true ones."""
    import torch

    rng = np.random.default_rng(44)
    Z, pos, _ = wl.random_molecule(500, 7e-9 * (500 / 20000) ** (1 / 3), rng)
    B = np.full(500, 20.0, np.float32)
    det = wl.Detector(48, 1.6e-3, 0.1, 0.2e-9)
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    ff = oracle.default_form_factors()
    _, I_true = oracle.sf_f64(q, Z, pos, atom_b=B)
    K = (1e3 / I_true.mean()) * np.ones_like(I_true)
    y = rng.poisson(K * I_true).astype(np.float32)
    pert = (pos + rng.normal(0, 0.05e-9, pos.shape)).astype(np.float32)

    # GPU: hand-written adjoint behind torch.autograd
    layout = to.SpeciesLayout(Z, gpu)
    hkl = torch.from_numpy(q).to(gpu)
    fft = dv.form_factor_table(device=gpu)
    P = torch.from_numpy(pert).to(gpu).requires_grad_(True)
    Bt = torch.from_numpy(B).to(gpu).requires_grad_(True)
    I = to.intensity(P, hkl, layout, fft, atom_b=Bt)
    loss = to.poisson_nll(I, torch.from_numpy(y).to(gpu), torch.from_numpy(K.astype(np.float32)).to(gpu))
    loss.backward()

    # CPU: dense FP64 autograd
    Pd = torch.from_numpy(pert.astype(np.float64)).requires_grad_(True)
    Bd = torch.from_numpy(B.astype(np.float64)).requires_grad_(True)
    Id = dense_torch_intensity(q, Z, Pd, Bd, ff, torch)
    lam = torch.from_numpy(K) * Id
    Ld = (lam - torch.from_numpy(y.astype(np.float64)) * torch.log(lam)).sum()
    Ld.backward()
    assert float(loss) == pytest.approx(float(Ld), rel=1e-5)
    assert rel_l2(P.grad.cpu().numpy(), Pd.grad.numpy()) < 1e-4
    assert rel_l2(Bt.grad.cpu().numpy(), Bd.grad.numpy()) < 1e-4


def nll_central_differences(q, F64, Z, pos, B, y, K, atoms, h=5e-14, dB=0.01, threads=0):
    """FP64 central differences of the Poisson NLL L = sum_p (K I_p - y_p log(K I_p))
    with respect to the positions and B factors of the listed atoms, at the
    FP32 inputs (pos, B) whose FP64 forward is F64. Moving one atom changes F by
    that atom's own term only, so L(x +- h) is evaluated exactly from F64 and
    the atom's term at the shifted inputs (oracle.sf_f64 of that one atom),
    with I+ - I- and log(I+/I-) formed directly to avoid cancellation. Steps
    are taken in FP32 (the step used is the FP64 difference of the two FP32
    inputs). Returns (fd_pos [len(atoms), 3], fd_b [len(atoms)])."""
    y = np.asarray(y, np.float64)

    def term(j, p, b):
        return oracle.sf_f64(q, Z[j:j + 1], p[None, :], atom_b=np.array([b], np.float32), threads=threads)[0]

    def dL(f0, j, p_plus, b_plus, p_minus, b_minus):
        fp, fm = term(j, p_plus, b_plus), term(j, p_minus, b_minus)
        ep, em = fp - f0, fm - f0
        dI = 2.0 * np.real(np.conj(F64) * (fp - fm)) + np.abs(ep) ** 2 - np.abs(em) ** 2
        Im = np.abs(F64 + em) ** 2
        return float(np.sum(K * dI - y * np.log1p(dI / Im)))

    fd_p = np.zeros((len(atoms), 3))
    fd_b = np.zeros(len(atoms))
    for k, j in enumerate(atoms):
        f0 = term(j, pos[j], B[j])
        for c in range(3):
            pp, pm = pos[j].copy(), pos[j].copy()
            pp[c] += np.float32(h)
            pm[c] -= np.float32(h)
            fd_p[k, c] = dL(f0, j, pp, B[j], pm, B[j]) / (float(pp[c]) - float(pm[c]))
        bp, bm = np.float32(B[j] + np.float32(dB)), np.float32(B[j] - np.float32(dB))
        fd_b[k] = dL(f0, j, pos[j], bp, pos[j], bm) / (float(bp) - float(bm))
    return fd_p, fd_b


def test_config4_full_workload_gradient(gpu):
    """Differentiation configuration at its full size: 20,000 C/N/O/S atoms
    (radius 7 nm), B = 20 A^2, 512 x 512 detector (75 um at 0.1 m, 0.2 nm).
    Counts y ~ Poisson(K I) at 1e3 mean photons per pixel for the true
    coordinates; the model is evaluated at coordinates perturbed by
    N(0, 0.05 nm). dL/dr and dL/dB of the GPU adjoint (behind autograd) on 64
    random atoms against the FP64 oracle gradient, and both against FP64
    central differences of the FP64 loss."""
    import torch

    det, Z, pos = wl.config4()
    n = len(Z)
    rng = np.random.default_rng(40)
    B = np.full(n, 20.0, np.float32)
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    layout = to.SpeciesLayout(Z, gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    Bt = torch.from_numpy(B).to(gpu)
    with torch.no_grad():
        I_true = to.intensity(torch.from_numpy(pos).to(gpu), hkl, layout, ff, atom_b=Bt).double().cpu().numpy()
    K = float(np.float32(1e3 / I_true.mean()))
    y = rng.poisson(K * I_true).astype(np.float32)
    pert = (pos + rng.normal(0, 0.05e-9, pos.shape)).astype(np.float32)

    P = torch.from_numpy(pert).to(gpu).requires_grad_(True)
    Bv = Bt.clone().requires_grad_(True)
    I = to.intensity(P, hkl, layout, ff, atom_b=Bv)
    loss = to.poisson_nll(I, torch.from_numpy(y).to(gpu), torch.full_like(I, K))
    loss.backward()
    gp_all, gb_all = P.grad.cpu().numpy(), Bv.grad.cpu().numpy()
    assert np.isfinite(gp_all).all() and np.isfinite(gb_all).all()

    F64, I64 = oracle.sf_f64(q, Z, pert, atom_b=B, threads=oracle_threads())
    L64 = float(np.sum(K * I64 - y * np.log(K * I64)))
    assert float(loss) == pytest.approx(L64, rel=1e-6)
    g = K - y.astype(np.float64) / I64
    sub = np.sort(rng.choice(n, 64, replace=False)).astype(np.int32)
    gp_ref, gb_ref = oracle.sf_grad_given_f_f64(q, F64, Z, pert, g, sub, atom_b=B, threads=oracle_threads())
    fd_p, fd_b = nll_central_differences(q, F64, Z, pert, B, y, K, sub, threads=oracle_threads())
    e_p, e_b = rel_l2(gp_all[sub], gp_ref), rel_l2(gb_all[sub], gb_ref)
    print(f"config 4 (20000 atoms, 512^2): GPU vs FP64 oracle gradient rel L2 pos {e_p:.2e} B {e_b:.2e}; "
          f"oracle vs central differences pos {rel_l2(gp_ref, fd_p):.2e} B {rel_l2(gb_ref, fd_b):.2e}")
    assert rel_l2(gp_ref, fd_p) < 1e-6 and rel_l2(gb_ref, fd_b) < 1e-6
    assert e_p < 1e-4 and e_b < 1e-4
    assert rel_l2(gp_all[sub], fd_p) < 1e-4 and rel_l2(gb_all[sub], fd_b) < 1e-4


def test_config4_n500_full_detector_vs_dense_fp64_torch(gpu):
    """The dense FP64 PyTorch reference at N = 500 on the configuration's full
    512 x 512 detector, evaluated in pixel chunks on the CPU (the loss is a sum
    over pixels, so the gradients of the chunks add up)."""
    import torch

    rng = np.random.default_rng(45)
    Z, pos, _ = wl.random_molecule(500, 7e-9 * (500 / 20000) ** (1 / 3), rng)
    B = np.full(500, 20.0, np.float32)
    det = wl.CFG4_DET
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    ff = oracle.default_form_factors()
    _, I_true = oracle.sf_f64(q, Z, pos, atom_b=B, threads=oracle_threads())
    K = float(np.float32(1e3 / I_true.mean()))
    y = rng.poisson(K * I_true).astype(np.float32)
    pert = (pos + rng.normal(0, 0.05e-9, pos.shape)).astype(np.float32)

    layout = to.SpeciesLayout(Z, gpu)
    hkl = torch.from_numpy(q).to(gpu)
    fft = dv.form_factor_table(device=gpu)
    P = torch.from_numpy(pert).to(gpu).requires_grad_(True)
    Bt = torch.from_numpy(B).to(gpu).requires_grad_(True)
    I = to.intensity(P, hkl, layout, fft, atom_b=Bt)
    loss = to.poisson_nll(I, torch.from_numpy(y).to(gpu), torch.full_like(I, K))
    loss.backward()

    Pd = torch.from_numpy(pert.astype(np.float64)).requires_grad_(True)
    Bd = torch.from_numpy(B.astype(np.float64)).requires_grad_(True)
    Ld = 0.0
    chunk = 16384
    for p0 in range(0, q.shape[0], chunk):
        Id = dense_torch_intensity(q[p0:p0 + chunk], Z, Pd, Bd, ff, torch)
        lam = K * Id
        part = (lam - torch.from_numpy(y[p0:p0 + chunk].astype(np.float64)) * torch.log(lam)).sum()
        part.backward()
        Ld += float(part)
    assert float(loss) == pytest.approx(Ld, rel=1e-6)
    assert rel_l2(P.grad.cpu().numpy(), Pd.grad.numpy()) < 1e-4
    assert rel_l2(Bt.grad.cpu().numpy(), Bd.grad.numpy()) < 1e-4


def test_poisson_nll_kernel(gpu):
    import torch

    rng = np.random.default_rng(9)
    I = rng.uniform(0.1, 5, 5000).astype(np.float32)
    y = rng.poisson(3.0, 5000).astype(np.float32)
    K = rng.uniform(0.5, 2, 5000).astype(np.float32)
    It = torch.from_numpy(I).to(gpu).requires_grad_(True)
    loss = to.poisson_nll(It, torch.from_numpy(y).to(gpu), torch.from_numpy(K).to(gpu))
    loss.backward()
    lam = K.astype(np.float64) * I
    ref = (lam - y * np.log(lam)).sum()
    assert float(loss) == pytest.approx(ref, rel=1e-12)
    np.testing.assert_allclose(It.grad.cpu().numpy(), K * (1 - y / lam), rtol=1e-6)
