"""Pins the proposal oracle (oracle/mlp.py): positional encoding closed
form, bf16 rounding, the loss-gradient closed form, float64 finite
differences of the network, bin targets and sampler invariants."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp as om


def test_positional_encoding_closed_form():
    p = torch.tensor([[0.25, -0.5, 0.125]], dtype=torch.float64)
    pe = om.positional_encoding(p)[0]
    assert pe.shape[0] == om.PE_DIM == 39
    assert torch.allclose(pe[:3], p[0])
    for k in range(6):
        a = (2 ** k) * math.pi * p[0]
        assert torch.allclose(pe[3 + 6 * k:6 + 6 * k], torch.sin(a))
        assert torch.allclose(pe[6 + 6 * k:9 + 6 * k], torch.cos(a))
    # sin(pi/4)*... spot values
    assert abs(float(pe[3]) - math.sin(math.pi / 4)) < 1e-15


def test_bf16_rounding_is_round_to_nearest_even():
    x = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.0], dtype=torch.float64)
    r = om.bf16_round(x)
    assert r.tolist() == [1.0, 1.0 + 2 ** -6, 1.0, -3.0]


def test_loss_gradient_closed_form_and_weights():
    g = torch.Generator().manual_seed(0)
    z = torch.randn((5, 64), generator=g, dtype=torch.float64).requires_grad_(True)
    q = torch.rand((5, 64), generator=g, dtype=torch.float64)
    q = q / q.sum(1, keepdim=True) * torch.rand((5, 1), generator=g, dtype=torch.float64)
    om.ray_loss(z, q).backward()
    assert torch.allclose(z.grad, om.ray_loss_grad(z.detach(), q))
    # weights and pass-through mass form a distribution
    w, T = om.proposal_weights(z.detach())
    assert torch.allclose(w.sum(1) + T, torch.ones(5, dtype=torch.float64))
    # the loss is minimised (gradient zero) when w reproduces the target
    zt = torch.log(torch.expm1(-torch.log1p(-(q / (1 - torch.cat([torch.zeros((5, 1), dtype=torch.float64), torch.cumsum(q, 1)[:, :-1]], 1))))))
    assert torch.allclose(om.ray_loss_grad(zt, q), torch.zeros_like(q), atol=1e-9)


def test_network_finite_differences():
    theta = om.init_params(1).double().requires_grad_(True)
    p = (torch.rand((6, 3), generator=torch.Generator().manual_seed(2), dtype=torch.float64) * 2 - 1)
    # without bf16 rounding the network is smooth almost everywhere: check
    # autograd against central differences on a few random directions
    f = lambda th: om.forward(th, p, emulate_bf16=False).sum()  # noqa: E731
    f(theta).backward()
    g = torch.Generator().manual_seed(3)
    for _ in range(4):
        v = torch.randn(theta.shape, generator=g, dtype=torch.float64)
        h = 1e-6
        fd = (f(theta.detach() + h * v) - f(theta.detach() - h * v)) / (2 * h)
        assert abs(float(fd) - float((theta.grad * v).sum())) < 1e-6 * (1 + abs(float(fd)))


def test_bin_targets_partition_the_weights():
    w = torch.rand((3, 10), dtype=torch.float64)
    rc = torch.tensor([10, 7, 0])
    q = om.bin_targets(w, rc)
    assert torch.allclose(q.sum(1), torch.stack([w[0].sum(), w[1, :7].sum(), torch.zeros((), dtype=torch.float64)]))
    # render sample j of 10 sits at (j+0.5)/10 of the segment: sample 4 in bin 28
    assert float(q[0, 28]) == float(w[0, 4])


def test_sampler_invariants():
    z = torch.full((1, 64), -60.0, dtype=torch.float64)
    t = om.fine_samples(z, torch.tensor([2.0]), K=34, eps=0.05)[0]
    assert t[0] == 0 and t[-1] == 2.0
    # a (near-)uniform pdf gives uniformly spaced interior samples
    assert torch.allclose(t[1:-1], (torch.arange(32, dtype=torch.float64) + 0.5) / 32 * 2.0,
                          atol=1e-6)
    # an opaque bin 10 takes the mass, and hides the opaque bin 40 behind it
    z = torch.full((1, 64), -50.0, dtype=torch.float64)
    z[0, 10] = 50.0
    z[0, 40] = 50.0
    t = om.fine_samples(z, torch.tensor([64.0]), K=32, eps=0.05)[0]
    inside = ((t >= 10) & (t < 11)).sum()
    assert int(inside) >= 27 and torch.all(t[1:] > t[:-1])


def test_sampler_sharpening_exponent():
    """Two surfaces of composited weight 0.2 (bin 5) and 0.6 (bin 20): the
    interior samples split 1:3 at gamma = 1 and 1:9 at gamma = 2."""
    def logit(d):
        return math.log(math.expm1(d))
    z = torch.full((1, 64), -60.0, dtype=torch.float64)
    z[0, 5] = logit(-math.log(1 - 0.2))
    z[0, 20] = logit(math.log(4.0))
    w, _ = om.proposal_weights(z)
    assert abs(float(w[0, 5]) - 0.2) < 1e-9 and abs(float(w[0, 20]) - 0.6) < 1e-9
    for gamma, n5 in ((1.0, 7), (2.0, 3)):
        t = om.fine_samples(z, torch.tensor([64.0]), K=32, eps=0.0, gamma=gamma)[0]
        assert int(((t >= 5) & (t < 6)).sum()) == n5
        assert int(((t >= 20) & (t < 21)).sum()) == 30 - n5


def test_composite_t_matches_list_compositing():
    from oracle import render as orc
    g = torch.Generator().manual_seed(4)
    sdf = torch.randn(9, generator=g, dtype=torch.float64) * 0.2
    rgb = torch.rand((8, 3), generator=g, dtype=torch.float64)
    ray = np.array([[9, 0, 8, 0]], dtype=np.int32)
    assert torch.allclose(om.composite_t(rgb, sdf, 6.0), orc.composite(rgb, sdf, ray, 6.0)[0])


def test_bin_centres_closed_form():
    """Bin b of a ray with c in-cube steps spans [b L/64, (b+1) L/64) of
    L = (c - 1) * interval from the first in-cube step, so its centre lies
    at (b + 0.5) / 64 * L along the unit direction.  With c = 65 (64 render
    samples) the centres coincide with the render samples of the ray
    generator, ori + dir * ((b + 0.5) * interval); bin_targets puts render
    sample b in bin b.  With c = 129 each centre lies midway between render
    samples 2b and 2b + 1, both of which bin_targets puts in bin b."""
    from oracle import raygen
    interval = 0.03
    g = torch.Generator().manual_seed(9)
    ori = torch.rand((2, 3), generator=g, dtype=torch.float64) * 0.4 - 0.2
    d = torch.randn((2, 3), generator=g, dtype=torch.float64)
    d = d / d.norm(dim=1, keepdim=True)
    b = torch.arange(64, dtype=torch.float64)
    p = om.bin_centres(ori, d, torch.tensor([65, 129]), interval)
    assert torch.allclose(p[0], ori[0] + d[0] * ((b + 0.5) * interval)[:, None], atol=1e-15)
    assert torch.allclose(p[1], ori[1] + d[1] * ((2 * b + 1) * interval)[:, None], atol=1e-15)
    # the ray generator's render sample j (cube coordinates) is the c = 65 bin centre j
    R = 64
    o32, d32 = ori[:1].numpy().astype(np.float32), d[:1].numpy().astype(np.float32)
    ray = np.array([[65, 0, 64, 0]], dtype=np.int32)
    _, _, rp, ri, _ = raygen.generate_ray_points(ray, o32, d32, np.float32(interval), R)
    ri = ri.astype(np.int64)
    cell = np.stack([ri % R, (ri // R) % R, ri // (R * R)], 1)
    pos = (2 * cell + 1 + rp.astype(np.float64)) / R - 1
    assert np.allclose(pos, p[0].numpy(), atol=2e-6)
    # bin_targets: one-hot weights at render sample j land in the bin whose centre they straddle
    eye = torch.eye(64, dtype=torch.float64)
    assert torch.equal(om.bin_targets(eye, torch.full((64,), 64)), eye)
    w128 = torch.zeros((128, 128), dtype=torch.float64)
    w128[torch.arange(128), torch.arange(128)] = 1.0
    q = om.bin_targets(w128, torch.full((128,), 128))
    assert torch.equal(q.argmax(1), torch.arange(128) // 2)
    # degenerate rays (c < 2) get their origin
    p0 = om.bin_centres(ori, d, torch.tensor([1, 0]), interval)
    assert torch.equal(p0, ori[:, None, :].expand(2, 64, 3))
