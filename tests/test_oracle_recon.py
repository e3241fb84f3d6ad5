"""Pins the oracle's reconstruction restatement (particle_reconstruction.hpp)
and the box reconstruction estimators (powerspec.hpp:396-519):

  - Psi(x) of a single Fourier mode against the closed form
    Psi_a(x) = 2 (G(k0)/b) (k0_a/k0^2) Im(A e^{i k0.x}) / V;
  - the TSC gather as the adjoint of the TSC assignment: the displacement of a
    particle equals norm * dV * sum_c rho_p(c) Psi(c), with rho_p the density
    calcField gives that particle alone (spills included);
  - the box reconstruction P_l against the plain box P_l (alpha = 0) and its
    quadratic dependence on alpha when the randoms coincide with the data.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

from .conftest import clustered_catalogue, uniform_catalogue


def test_psi_single_mode_closed_form():
    n, L, b1, RG = 16, 400.0, 2.0, 15.0
    p = po.param(n, L, po.TSC)
    dk = np.zeros((n, n, n), complex)
    A = 1.5 - 0.7j
    m0 = np.array([2, -1, 3])
    dk[tuple(m0 % n)] = A
    dk[tuple(-m0 % n)] = np.conj(A)
    psi = po.recon_psi(p, dk, b1, RG)
    k0 = m0 * 2 * np.pi / L
    kk = np.sum(k0 ** 2)
    G = np.exp(-(np.sqrt(kk) * RG) ** 2 / 2)
    x = np.stack(np.meshgrid(*[np.arange(n) * L / n] * 3, indexing="ij"), -1)
    phase = np.exp(1j * (x @ k0))
    for a in range(3):
        expect = 2 * (G / b1) * (k0[a] / kk) * np.imag(A * phase) / L ** 3
        assert np.allclose(psi[a].real, expect, rtol=0, atol=1e-13 * np.abs(expect).max())
        assert np.abs(psi[a].imag).max() <= 1e-13 * np.abs(expect).max()


def test_gather_is_adjoint_of_assignment():
    n, L, norm = 8, 200.0, 3.7
    p = po.param(n, L, po.TSC)
    rng = np.random.default_rng(2)
    psi = rng.normal(size=(3, n, n, n)) + 0j
    dx = L / n
    pos = rng.random((25, 3)) * L
    pos[0] = [3.3 * dx, 4.2 * dx, (n - 0.3) * dx]  # z index n: spills into the next row
    pos[1] = [(n - 0.2) * dx, 2.0 * dx, 2.0 * dx]  # x index n: dropped part
    moved = po.recon_gather(p, psi, norm, pos)
    dV = dx ** 3
    for q in range(len(pos)):
        rho = po.calc_field(p, pos[q:q + 1]).real
        expect = pos[q] + norm * dV * np.array([np.sum(rho * psi[a].real) for a in range(3)])
        assert np.allclose(moved[q], expect, rtol=1e-13, atol=1e-12 * norm * np.abs(psi).max())


def test_box_recon_pk_against_plain_pk_and_alpha():
    n, L = 24, 1000.0
    posD = clustered_catalogue(6000, L, seed=3)
    p = po.param(n, L, po.TSC, kmin=0.01, kmax=0.15, n_kbin=8)
    ref, nm_ref = po.box_pk(p, posD)
    got, nm = po.box_recon_2pt(p, posD, posD, 0.0)
    assert np.array_equal(nm, nm_ref)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    # randoms = data: P(alpha) = (1 - alpha)^2 A - (1 + alpha^2) B with A - B = P(0), -2 B = P(1)
    B = -po.box_recon_2pt(p, posD, posD, 1.0)[0] / 2
    A = got + B
    for alpha in (0.3, 0.8):
        pa, _ = po.box_recon_2pt(p, posD, posD, alpha)
        assert np.allclose(pa, (1 - alpha) ** 2 * A - (1 + alpha ** 2) * B, rtol=1e-9, atol=1e-9 * np.abs(A).max())


def test_recon_displace_moves_particles_by_the_gathered_field():
    """The full displacement equals Psi from the corrected spectrum of D - alpha R, gathered."""
    n, L, b1, RG = 16, 600.0, 2.0, 20.0
    p = po.param(n, L, po.TSC)
    posD = clustered_catalogue(3000, L, seed=4)
    posR = uniform_catalogue(9000, L, seed=5)
    alpha = float(3000 // 9000)  # Q4: integer division, 0 here
    dD, dR = po.recon_displace(p, posD, None, posR, None, alpha, L ** 3 / 3000, b1, RG)
    f = po.calc_field(p, posD) - alpha * po.calc_field(p, posR)
    dk = po.fourier_transform(p, f)
    kk = 2 * np.pi / L * np.array([i if i < n // 2 else i - n for i in range(n)])
    W = np.array([[[po.assignment_function(p, [a, b, c]) for c in kk] for b in kk] for a in kk])
    psi = po.recon_psi(p, dk / W, b1, RG)
    assert np.allclose(dD, po.recon_gather(p, psi, L ** 3 / 3000, posD), rtol=1e-13, atol=1e-10)
    assert np.allclose(dR, po.recon_gather(p, psi, L ** 3 / 3000, posR), rtol=1e-13, atol=1e-10)
    assert np.abs(dD - posD).max() > 0.1  # it moved
