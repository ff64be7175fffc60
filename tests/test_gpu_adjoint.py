"""The hand-written adjoint of the run (prop/adjoint.py) on the HIP path:

* equal to torch.autograd through the differentiable restatement
  (oracle/adjoint_oracle.py, itself pinned to central differences of the numpy
  oracle) on small systems, to 1e-8 relative;
* equal to central finite differences with common random numbers on the
  96-orbital, 16-occupied, 480-vector, 512-walker workload (dt 0.01, 50 steps
  in blocks of 10 = one checkpoint per 10 steps, seed 3, V = 0.01 sym(N(0,1)),
  lambda = +-1e-4), to the 1% the target names -- in practice far closer.
"""
import numpy as np
import pytest

from ad_afqmc_prototype_amd import synthetic as syn
from oracle import adjoint_oracle as ado

from .test_adjoint import _tiny

pytestmark = pytest.mark.gpu


def _run_from_spec(spec, device):
    import torch
    from ad_afqmc_prototype_amd.prop import adjoint as adj
    return adj.adjoint_run(trial=torch.tensor(spec.trial, device=device), h0=spec.h0, dt=spec.dt,
                           n_walkers=spec.n_walkers, n_blocks=spec.n_blocks,
                           n_prop_steps=spec.n_prop_steps,
                           fields=torch.tensor(spec.fields, device=device),
                           zetas=torch.tensor(spec.zetas, device=device))


@pytest.mark.parametrize("kw", [
    dict(seed=11, n=24, o=6, g=48, nw=32, n_blocks=2, n_prop=3, scale=0.25),
    dict(seed=3, n=20, o=4, g=40, nw=48, n_blocks=3, n_prop=4, scale=0.6),
    dict(seed=4, n=16, o=3, g=32, nw=24, n_blocks=2, n_prop=5, scale=1.5, dt=0.02),
])
def test_adjoint_matches_autograd_restatement(gpu_device, kw):
    import torch
    from ad_afqmc_prototype_amd.prop import adjoint as adj
    print(f"adjoint vs autograd restatement {kw}", flush=True)
    h0, h1, chol, c, spec = _tiny(**kw)
    j_ref, gh_ref, gl_ref = ado.gradients(h0, h1, chol, spec)
    run = _run_from_spec(spec, gpu_device)
    out = adj.energy_gradients(torch.tensor(h1, device=gpu_device), torch.tensor(chol, device=gpu_device),
                               run)
    assert abs(out["energy"] - j_ref) <= 1e-11 * abs(j_ref)
    gh = out["d_h1"].cpu().numpy()
    gl = out["d_chol"].cpu().numpy()
    eh = np.abs(gh - gh_ref).max() / np.abs(gh_ref).max()
    el = np.abs(gl - gl_ref).max() / np.abs(gl_ref).max()
    assert eh < 1e-8 and el < 1e-8, (eh, el)


@pytest.mark.parametrize("n,o", [(24, 6), (96, 16), (120, 20), (128, 32)])
def test_trotter_adjoint_kernel_matches_autograd(gpu_device, n, o):
    """afq_trotter_adjoint against torch.autograd through a complex128
    restatement of the Trotter step (CPU), for a general real E (so E^T is
    exercised) and a complex-symmetric V: w_bar, V_bar and E_bar to 1e-12
    relative.  (n, o) = (120, 20), (128, 32): two column tiles and V k-blocks
    kept in LDS."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    nw, K, dt = 5, 6, 0.01
    g = torch.Generator().manual_seed(n + o)
    cplx = lambda *sh: torch.complex(torch.randn(*sh, generator=g, dtype=torch.float64),  # noqa: E731
                                     torch.randn(*sh, generator=g, dtype=torch.float64))
    w = cplx(nw, n, o)
    v = cplx(nw, n, n)
    v = 0.5 * (v + v.transpose(1, 2)) * (0.3 / n ** 0.5 / dt ** 0.5)
    e = torch.eye(n, dtype=torch.float64) + 0.01 * torch.randn(n, n, generator=g, dtype=torch.float64)
    wb = cplx(nw, n, o)
    # reference: J = Re sum conj(wb) w'
    wl, vl, el = w.clone().requires_grad_(), v.clone().requires_grad_(), e.clone().requires_grad_()
    a = 1.0j * dt ** 0.5
    t = el.to(torch.complex128) @ wl
    u = t
    for k in range(1, K + 1):
        t = (a / k) * (vl @ t)
        u = u + t
    wp = el.to(torch.complex128) @ u
    torch.sum(wb.conj() * wp).real.backward()
    ldv = kernels.ldv_for(n)
    dev = gpu_device
    e_pad = torch.zeros((n, ldv), dtype=torch.float64)
    e_pad[:, :n] = e
    e_pad_t = torch.zeros((n, ldv), dtype=torch.float64)
    e_pad_t[:, :n] = e.T
    vp = torch.zeros((2, nw, n, ldv), dtype=torch.float64)
    vp[0, :, :, :n], vp[1, :, :, :n] = v.real, v.imag
    tc = (o + 15) // 16 * 16
    ws = torch.empty((2 * (2 * K + 2) * nw * n * tc,), dtype=torch.float64, device=dev)
    wd, wbd = w.to(dev), wb.to(dev)
    w_bar = torch.empty_like(wd)
    v_bar = torch.empty((2, nw, n, n), dtype=torch.float64, device=dev)
    e_bar = torch.empty((nw, n, n), dtype=torch.float64, device=dev)
    e_pad, e_pad_t, vp = e_pad.to(dev), e_pad_t.to(dev), vp.to(dev)
    kernels._call("afq_trotter_adjoint", wd, nw, n, o, ldv, K, dt ** 0.5, e_pad.data_ptr(), e_pad_t.data_ptr(),
                  vp.data_ptr(), wd.data_ptr(), wbd.data_ptr(), w_bar.data_ptr(), v_bar.data_ptr(),
                  e_bar.data_ptr(), ws.data_ptr(), kernels._stream(wd))
    torch.cuda.synchronize()
    rel = lambda x, y: float((x - y).abs().max() / y.abs().max())  # noqa: E731
    assert rel(w_bar.cpu(), wl.grad) < 1e-12
    assert rel(torch.complex(v_bar[0], v_bar[1]).cpu(), vl.grad) < 1e-12
    assert rel(e_bar.sum(0).cpu(), el.grad) < 1e-12


@pytest.mark.parametrize("n,o,parts", [(24, 6, "both"), (96, 16, "g"), (96, 16, "ld"), (120, 32, "both"),
                                        (128, 27, "both")])
def test_greens_adjoint_kernel_matches_autograd(gpu_device, n, o, parts):
    """afq_greens_adjoint against torch.autograd (CPU, complex128) of
    G = (w m^-1)^T and log det m, m = C^T w, for J = Re sum conj(G_bar) G +
    Re sum conj(ld_bar) ld: w_bar to 1e-11 relative, written and accumulated."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    nw = 6
    g = torch.Generator().manual_seed(7 * n + o)
    cplx = lambda *sh: torch.complex(torch.randn(*sh, generator=g, dtype=torch.float64),  # noqa: E731
                                     torch.randn(*sh, generator=g, dtype=torch.float64))
    c = torch.linalg.qr(torch.randn(n, o, generator=g, dtype=torch.float64))[0].contiguous()
    w = (c.to(torch.complex128)[None] + 0.3 * cplx(nw, n, o)).contiguous()
    gb = cplx(nw, o, n) if parts in ("g", "both") else None
    lb = cplx(nw) if parts in ("ld", "both") else None
    wl = w.clone().requires_grad_()
    m = c.T.to(torch.complex128) @ wl
    j = 0.0
    if gb is not None:
        gmat = (wl @ torch.linalg.inv(m)).transpose(1, 2)
        j = j + torch.sum(gb.conj() * gmat).real
    if lb is not None:
        j = j + torch.sum(lb.conj() * torch.linalg.slogdet(m)[1]).real
        j = j + torch.sum((lb.conj() * torch.log(torch.linalg.slogdet(m)[0]))).real
    j.backward()
    dev = gpu_device
    wd, cd = w.to(dev), c.to(dev).contiguous()
    gbd = gb.to(dev) if gb is not None else None
    lbd = lb.to(dev) if lb is not None else None
    out = torch.empty_like(wd)
    kernels._call("afq_greens_adjoint", wd, nw, n, o, wd.data_ptr(), cd.data_ptr(),
                  gbd.data_ptr() if gbd is not None else None, lbd.data_ptr() if lbd is not None else None,
                  out.data_ptr(), 0, kernels._stream(wd))
    base = cplx(nw, n, o).to(dev)
    acc = base.clone()
    kernels._call("afq_greens_adjoint", wd, nw, n, o, wd.data_ptr(), cd.data_ptr(),
                  gbd.data_ptr() if gbd is not None else None, lbd.data_ptr() if lbd is not None else None,
                  acc.data_ptr(), 1, kernels._stream(wd))
    torch.cuda.synchronize()
    ref = wl.grad
    err = float((out.cpu() - ref).abs().max() / ref.abs().max())
    assert err < 1e-11, err
    assert torch.allclose(acc - base, out, rtol=0, atol=1e-12 * float(out.abs().max()))


def _workload(device, n_blocks=5, n_prop=10):
    import torch
    from ad_afqmc_prototype_amd.prop import adjoint as adj
    cfg = syn.WORKLOADS["n96_o16_g480_w512_rhf"]
    s = syn.make_system(cfg["seed"], cfg["norb"], cfg["nocc"], cfg["naux"], chol_scale=cfg["chol_scale"])
    fields, zetas = adj.device_random_numbers(cfg["seed"], n_blocks, n_prop, cfg["n_walkers"], cfg["naux"],
                                              device)
    run = adj.adjoint_run(trial=torch.tensor(s.mo_up, device=device), h0=float(s.h0), dt=cfg["dt"],
                          n_walkers=cfg["n_walkers"], n_blocks=n_blocks, n_prop_steps=n_prop,
                          fields=fields, zetas=zetas)
    rng = np.random.default_rng(cfg["seed"] + 1000)
    m = rng.standard_normal((cfg["norb"], cfg["norb"]))
    v = 0.01 * 0.5 * (m + m.T)
    return run, torch.tensor(s.h1, device=device), torch.tensor(s.chol, device=device), \
        torch.tensor(v, device=device)


def _central(h1, chol, run, dh1=None, dchol=None, eps=1e-4):
    from ad_afqmc_prototype_amd.prop import adjoint as adj
    jp, ebp, wbp, idxp = adj.run_blocks(h1 + eps * dh1 if dh1 is not None else h1,
                                        chol + eps * dchol if dchol is not None else chol, run)
    jm, ebm, wbm, idxm = adj.run_blocks(h1 - eps * dh1 if dh1 is not None else h1,
                                        chol - eps * dchol if dchol is not None else chol, run)
    same = all(bool((a == b).all()) for a, b in zip(idxp, idxm))
    return (jp - jm) / (2.0 * eps), same


def test_adjoint_matches_finite_differences_96_orbitals(gpu_device):
    import torch
    from ad_afqmc_prototype_amd.prop import adjoint as adj
    print("adjoint vs central differences, 96 orbitals / 512 walkers / 50 steps", flush=True)
    run, h1, chol, v = _workload(gpu_device)
    out = adj.energy_gradients(h1, chol, run, v)
    # the run is live: the block energies move and no walker population collapses
    j, eb, wb, idx = adj.run_blocks(h1, chol, run)
    assert abs(j - out["energy"]) <= 1e-12 * abs(j)
    assert float(wb.min()) > 0.5 * run.n_walkers
    assert float(eb.max() - eb.min()) > 1e-3
    d_fd, same = _central(h1, chol, run, dh1=v, eps=1e-4)
    assert same, "a comb pick differs between lambda = +-1e-4: the difference quotient is not valid"
    d_ad = out["d_lambda"]
    print(f"dE/dlambda adjoint {d_ad:.12e} central {d_fd:.12e} rel {abs(d_ad - d_fd) / abs(d_fd):.2e}",
          flush=True)
    assert abs(d_ad - d_fd) <= 1e-2 * abs(d_fd)
    assert abs(d_ad - d_fd) <= 1e-5 * abs(d_fd)  # what the double-precision run actually gives
    # d<E>/dL along a random symmetric direction of the Cholesky tensor's own scale
    gen = torch.Generator(device="cpu").manual_seed(7)
    a = torch.randn(tuple(chol.shape), dtype=torch.float64, generator=gen).to(gpu_device)
    d = 0.5 * (a + a.transpose(1, 2)) * float(chol.abs().mean())
    dl_ad = float(torch.sum(out["d_chol"] * d))
    dl_fd, same = _central(h1, chol, run, dchol=d, eps=1e-4)
    assert same
    print(f"dE/dL.D adjoint {dl_ad:.12e} central {dl_fd:.12e} rel {abs(dl_ad - dl_fd) / abs(dl_fd):.2e}",
          flush=True)
    assert abs(dl_ad - dl_fd) <= 1e-2 * abs(dl_fd)
    assert abs(dl_ad - dl_fd) <= 1e-5 * abs(dl_fd)
