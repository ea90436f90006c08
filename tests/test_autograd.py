"""bgy3d_amd.autograd without a GPU: the parameter <-> site mapping, the
Function's forward/backward wiring (torch.autograd.gradcheck with the solve
replaced by a closed-form mu whose gradient is known), and the loud failure
without a device."""
import numpy as np
import pytest
import torch

from bgy3d_amd import autograd as A
from bgy3d_amd import rism3d as R


def test_params_round_trip():
    sites = [("O", (0.1, -0.2, 0.3), 3.1, 0.15, -0.8), ("H", (1.0, 0.0, 0.0), 0.4, 0.05, 0.4)]
    names, p = A.params_from_sites(sites)
    assert names == ["O", "H"]
    assert p[0].tolist() == [0.1, -0.2, 0.3, -0.8, 3.1, 0.15]
    assert A.sites_from_params(names, p) == [(n, tuple(x), s, e, q) for n, x, s, e, q in sites]
    with pytest.raises(ValueError):
        A.sites_from_params(names, p[:, :5])


def test_function_wiring(monkeypatch):
    """A fake solve: mu = sum_a (q_a^2 sigma_a + eps_a |x_a|^2) with its
    exact gradient as 'free-energy-dsite'; gradcheck then tests that the
    Function passes the adjoint columns back in the params order."""
    def fake(sites, solvent, settings, chi=None, adjoint=False):
        names, p = A.params_from_sites(sites)
        x, q, s, e = p[:, :3], p[:, 3], p[:, 4], p[:, 5]
        mu = float(np.sum(q * q * s + e * np.sum(x * x, axis=1)))
        out = {"free-energy": mu, "status": "OK"}
        if adjoint:
            d = np.zeros_like(p)
            d[:, :3] = 2 * e[:, None] * x
            d[:, 3] = 2 * q * s
            d[:, 4] = q * q
            d[:, 5] = np.sum(x * x, axis=1)
            out["free-energy-dsite"] = d
        return out
    monkeypatch.setattr(R, "hnc3d_run_solute", fake)
    rng = np.random.default_rng(0)
    p = torch.tensor(rng.normal(size=(3, 6)), dtype=torch.float64, requires_grad=True)
    f = lambda t: A.solute_mu(t, ["a", "b", "c"], "water", {"N": 32})
    assert torch.autograd.gradcheck(f, (p,), eps=1e-6, atol=1e-8)
    with torch.no_grad():  # no gradient requested: no adjoint asked for
        assert float(f(p)) == pytest.approx(float(f(p.detach())))


def test_settings_reach_the_solve(monkeypatch):
    """The solve's settings pass through unchanged, the Ng path ('no-renorm',
    whose adjoint the library forms as on the tau path) and mixed precision
    included, and the adjoint is asked for only when the parameters need a
    gradient."""
    seen = []

    def fake(sites, solvent, settings, chi=None, adjoint=False):
        seen.append((dict(settings), adjoint))
        out = {"free-energy": 1.0, "status": "OK"}
        if adjoint:
            out["free-energy-dsite"] = np.zeros((len(sites), 6))
        return out
    monkeypatch.setattr(R, "hnc3d_run_solute", fake)
    st = {"no-renorm": True, "precision": "mixed", "N": 32}
    p = torch.zeros((1, 6), dtype=torch.float64, requires_grad=True)
    A.solute_mu(p, ["a"], "water", st).backward()
    A.solute_mu(p.detach(), ["a"], "water", st)
    assert seen == [(st, True), (st, False)]


def _gpu_present():
    return torch.cuda.is_available()


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU error path")
def test_no_device_fails_loudly():
    chi = np.tile(np.eye(3)[:, :, None], (1, 1, 1536))
    p = torch.tensor([[0.0, 0.0, 0.0, 0.0, 3.0, 0.1]], dtype=torch.float64, requires_grad=True)
    with pytest.raises(R.RismError) as e:
        A.solute_mu(p, ["X"], "water, PR-SPC/E", {"L": 8, "N": 32, "rmax": 40.0, "nrad": 1536},
                    chi)
    assert e.value.status == R.NO_DEVICE
