"""Generalized (spin-orbital) walkers in the restricted basis with an RHF trial
on the CPU checker (meas/rhf.py:108-130, trial/rhf.py:47-53,
prop/chol_afqmc_ops.py:118-212, walkers.py:19-49):

* force bias = d/dx log <T|(1 + x L_g)|w> at 0 with L_g acting on both spin
  blocks (the meas/auto.py:60-83 principle, which the reference's own
  tests/test_auto.py:136-164 checks the generalized kernel against);
* a spin-collinear walker [[wu, 0], [0, wd]] reproduces the unrestricted
  overlap, force bias and Trotter step exactly;
* no energy kernel (KeyError, meas/rhf.py:169-174).
"""
import numpy as np
import pytest

from oracle import afqmc_oracle as orc
from ad_afqmc_prototype_amd import synthetic as syn


def _setup(seed=3, n=7, o=2, g=9):
    h0, h1, chol = syn.make_hamiltonian(seed, n, g, chol_scale=0.4)
    c = np.linalg.eigh(h1)[1][:, :o].copy()
    return h0, h1, chol, (c, c)


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_force_bias_is_the_overlap_derivative():
    h0, h1, chol, trial = _setup()
    mctx = orc.build_meas_ctx(h1, chol, trial)
    rng = np.random.default_rng(0)
    n, o = trial[0].shape
    w = _rand(rng, 3, 2 * n, 2 * o)
    fb = orc.force_bias(w, trial, mctx, "generalized")
    for i in range(3):
        ref = orc.force_bias_fd(w[i], trial, chol, "generalized")
        np.testing.assert_allclose(fb[i], ref, rtol=1e-8, atol=1e-9)


def test_spin_collinear_walker_reduces_to_unrestricted():
    h0, h1, chol, trial = _setup()
    mctx = orc.build_meas_ctx(h1, chol, trial)
    rng = np.random.default_rng(1)
    n, o = trial[0].shape
    wu, wd = _rand(rng, 4, n, o), _rand(rng, 4, n, o)
    wg = np.zeros((4, 2 * n, 2 * o), dtype=np.complex128)
    wg[:, :n, :o], wg[:, n:, o:] = wu, wd
    np.testing.assert_allclose(orc.overlap(wg, trial, "generalized"),
                               orc.overlap((wu, wd), trial, "unrestricted"), rtol=1e-12)
    np.testing.assert_allclose(orc.force_bias(wg, trial, mctx, "generalized"),
                               orc.force_bias((wu, wd), trial, mctx, "unrestricted"), rtol=1e-11, atol=1e-12)
    pctx = orc.build_prop_ctx(h0, h1, chol, orc.get_rdm1(trial), 0.01)
    x = _rand(rng, 4, chol.shape[0])
    tg = orc.apply_trotter(wg, x, pctx, chol, 6, "generalized")
    tu, td = orc.apply_trotter((wu, wd), x, pctx, chol, 6, "unrestricted")
    np.testing.assert_allclose(tg[:, :n, :o], tu, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(tg[:, n:, o:], td, rtol=1e-12, atol=1e-14)
    assert np.abs(tg[:, :n, o:]).max() == 0.0 and np.abs(tg[:, n:, :o]).max() == 0.0


def test_init_walkers_and_no_energy():
    h0, h1, chol, trial = _setup()
    n, o = trial[0].shape
    w = orc.init_walkers(orc.get_rdm1(trial), o, 5, "generalized")
    assert w.shape == (5, 2 * n, 2 * o)
    assert np.abs(w[:, :n, o:]).max() == 0.0 and np.abs(w[:, n:, :o]).max() == 0.0
    # the natural orbitals span the trial's occupied space in each spin block
    np.testing.assert_allclose(abs(orc.overlap(w, trial, "generalized")), 1.0, rtol=1e-12)
    with pytest.raises(KeyError):
        orc.energy(w, h0, trial, orc.build_meas_ctx(h1, chol, trial), "generalized")
