"""The torch.autograd wrapper of the shape gradient (kigpu.shape.shape_loss)
on the CPU, with the device computation replaced by a stand-in of known
value and gradient: the forward pass calls it once, backward scales its
gradient by the incoming one, dtype and device follow the input, and an
optimizer step moves the coefficients down the gradient.  The device path
itself is checked against the dense complex-step gradient in
test_gpu_adjoint.py."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from kigpu import shape  # noqa: E402


@pytest.fixture
def stand_in(monkeypatch):
    calls = []

    def fake(ctx, star, c, targets, **kw):
        calls.append((np.array(c), kw))
        return float(np.sum(np.arange(1, c.size + 1) * c ** 2)), 2 * np.arange(1, c.size + 1) * c, {}
    monkeypatch.setattr(shape, "loss_and_gradient", fake)
    return calls


def test_backward_scales_the_kept_gradient(stand_in):
    c = torch.tensor([0.1, -0.2, 0.3, 0.05], dtype=torch.float64, requires_grad=True)
    L = shape.shape_loss(None, None, c, None, id_tol=1e-8)
    assert L.dtype == torch.float64 and L.shape == ()
    (3.0 * L).backward()
    w = np.arange(1, 5)
    assert np.allclose(L.item(), np.sum(w * c.detach().numpy() ** 2), rtol=0, atol=1e-15)
    assert np.allclose(c.grad.numpy(), 3.0 * 2 * w * c.detach().numpy(), rtol=0, atol=1e-15)
    assert len(stand_in) == 1 and stand_in[0][1] == {"id_tol": 1e-8}


def test_gradient_step(stand_in):
    c = torch.tensor([0.2, 0.2, -0.1, 0.3], dtype=torch.float64, requires_grad=True)
    opt = torch.optim.SGD([c], lr=0.05)
    losses = []
    for _ in range(5):
        opt.zero_grad()
        L = shape.shape_loss(None, None, c, None)
        L.backward()
        opt.step()
        losses.append(L.item())
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert len(stand_in) == 5
