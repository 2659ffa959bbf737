"""Float64 restatement of the proposal network and the fine sampler.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The proposal has
no counterpart in the reference; this module is its specification.

Network (per point p in the normalised cube):
  x0  = bf16( PE(p) )                       PE: 39 features, padded to 64
  x1  = bf16( relu(W1 x0 + b1) )            64
  x2  = bf16( relu(W2 x1 + b2) )            64
  h3  = relu(W3 x2 + b3)                    64, kept in float32
  z   = w4 . h3 + b4                        occupancy logit
All weight matrices are used at their bfloat16 values; products of bf16
operands are exact in float32 and sums are accumulated in float32 on the
GPU (float64 here).  Gradients treat the bf16 roundings as identity
(straight-through), i.e. they are the exact gradient of the computed
forward.

PE(p) = [p_x, p_y, p_z] + for k in 0..5:
        [sin(2^k pi p_x), sin(.. p_y), sin(.. p_z), cos(2^k pi p_x), cos(.. p_y), cos(.. p_z)]

Bins: a ray whose plain cut has c >= 2 in-cube steps spans
t in [0, L], L = (c - 1) * interval, from its first in-cube step; bin b is
[b L/64, (b+1) L/64) and the network is evaluated at the bin centres.

Proposal weights: the logit is read as a density d_b = softplus(z_b) of
bin b, composited front to back like the renderer:
a_b = 1 - exp(-d_b), T_b = exp(-sum_{c<b} d_c), w_b = T_b a_b and the
pass-through mass T_end = exp(-sum_c d_c).  Occluded bins (behind an
opaque one) therefore get no mass, which a per-ray softmax cannot express.

Loss per ray: the categorical log-likelihood of the plain renderer's
weights over the 64 bins plus the pass-through outcome,
  -sum_b q_b log w_b - (1 - sum_b q_b) log T_end,
with q_b the per-bin sums of the plain compositing weights.  Closed form:
  dL/dz_b = sigmoid(z_b) [ (1 - sum_{c<=b} q_c) - q_b / expm1(d_b) ].

Sampler: pdf_b = (1 - eps) w_b^g / sum w^g + eps / 64 with the sharpening
exponent g (g > 1 concentrates the samples where the proposal's weights
peak, compensating the smoothness of a small network); fine sdf samples are
t = 0, t = L and the inverse CDF of the piecewise-constant pdf at
u_k = (k + 0.5) / (K - 2) for k < K - 2, merged in increasing order.

Pinned by (tests/test_oracle_mlp.py): PE closed form, float64 finite
differences of the network, the loss gradient's closed form and its zero at
the target,
and sampler invariants (monotone, endpoints, bins with all the mass get
the interior samples, a uniform pdf gives uniform spacing).
"""
from __future__ import annotations

import math

import numpy as np
import torch

N_FREQ = 6
PE_DIM = 3 + 6 * N_FREQ          # 39
K_IN = 64                         # padded input width
HIDDEN = 64
NBINS = 64

# parameter vector layout (float32 master copy): W1[64,39] b1[64] W2[64,64]
# b2[64] W3[64,64] b3[64] w4[64] b4[1]
SHAPES = [("W1", (HIDDEN, PE_DIM)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)),
          ("b2", (HIDDEN,)), ("W3", (HIDDEN, HIDDEN)), ("b3", (HIDDEN,)),
          ("w4", (HIDDEN,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned in x's dtype."""
    return x.to(torch.float32).to(torch.bfloat16).to(x.dtype)


def _st(x):
    """Straight-through bf16 rounding (forward rounds, gradient passes)."""
    return x + (bf16_round(x) - x).detach()


def positional_encoding(p):
    """p [n,3] -> [n,39] (canonical order above)."""
    feats = [p]
    for k in range(N_FREQ):
        a = (2.0 ** k) * math.pi * p
        feats.append(torch.sin(a))
        feats.append(torch.cos(a))
    return torch.cat(feats, dim=1)


def init_params(seed=0):
    """He-uniform init (float32 master vector)."""
    g = torch.Generator().manual_seed(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("b"):
            parts.append(torch.zeros(shp))
        else:
            fan_in = shp[-1] if len(shp) == 2 else HIDDEN
            bound = math.sqrt(6.0 / fan_in)
            parts.append((torch.rand(shp, generator=g) * 2 - 1) * bound)
    return torch.cat([t.reshape(-1) for t in parts]).to(torch.float32)


def forward(theta, p, emulate_bf16=True):
    """Logits z [n] for points p [n,3]; theta float vector (any dtype)."""
    P = unpack(theta)
    r = _st if emulate_bf16 else (lambda x: x)
    W1, W2, W3, w4 = (bf16_round(P[k]) if emulate_bf16 else P[k] for k in ("W1", "W2", "W3", "w4"))
    x0 = r(positional_encoding(p))
    x1 = r(torch.relu(x0 @ W1.t() + P["b1"]))
    x2 = r(torch.relu(x1 @ W2.t() + P["b2"]))
    h3 = torch.relu(x2 @ W3.t() + P["b3"])
    return h3 @ w4 + P["b4"]


def bin_centres(ori, dirs, count, interval, nbins=NBINS):
    """Points at the bin centres of every ray: [n_rays, nbins, 3]; rays with
    count < 2 get their origin (their logits are unused)."""
    L = torch.clamp(count.to(ori.dtype) - 1, min=0) * interval
    t = (torch.arange(nbins, dtype=ori.dtype) + 0.5) / nbins
    return ori[:, None, :] + dirs[:, None, :] * (t[None, :, None] * L[:, None, None])


def proposal_weights(z):
    """(w [n, nbins], T_end [n]) of the density reading of the logits."""
    d = torch.nn.functional.softplus(z)
    S = torch.cumsum(d, dim=1)
    T = torch.exp(-(S - d))
    a = -torch.expm1(-d)
    return T * a, torch.exp(-S[:, -1])


def ray_loss(z, q):
    """sum over rays of -sum_b q_b log w_b - (1 - sum q) log T_end."""
    d = torch.nn.functional.softplus(z)
    S = torch.cumsum(d, dim=1)
    logw = -(S - d) + torch.log(-torch.expm1(-d))
    return (-(q * logw).sum(1) + (1 - q.sum(1)) * S[:, -1]).sum()


def ray_loss_grad(z, q):
    """Closed form dL/dz_b = sigmoid(z_b) [(1 - sum_{c<=b} q_c) - q_b / expm1(d_b)]."""
    d = torch.nn.functional.softplus(z)
    return torch.sigmoid(z) * ((1 - torch.cumsum(q, dim=1)) - q / torch.expm1(d))


def bin_targets(w, count_render, nbins=NBINS):
    """Per-bin sums of compositing weights.  w [n_rays, Lmax] (weights of the
    render samples in order), count_render [n_rays]; render sample j of a ray
    with c render samples lies at t = (j + 0.5) * interval, i.e. in bin
    floor((j + 0.5) / c * nbins)."""
    n, Lm = w.shape
    j = torch.arange(Lm, dtype=torch.float64)
    c = count_render.to(torch.float64).clamp_min(1)
    b = torch.floor((j[None, :] + 0.5) / c[:, None] * nbins).clamp(max=nbins - 1).long()
    valid = j[None, :] < count_render[:, None]
    q = torch.zeros((n, nbins), dtype=w.dtype)
    q.scatter_add_(1, torch.where(valid, b, 0), torch.where(valid, w, torch.zeros_like(w)))
    return q


def fine_samples(z, L, K=32, eps=0.05, gamma=1.0):
    """Inverse-CDF fine sample positions t [n_rays, K] from logits z
    [n_rays, nbins] on [0, L]; the pdf is the proposal weights raised to
    ``gamma``, normalised, mixed with eps of uniform."""
    n, nb = z.shape
    w, _ = proposal_weights(z.to(torch.float64))
    w = w ** gamma
    tot = w.sum(1, keepdim=True)
    wn = torch.where(tot > 1e-30, w / tot.clamp_min(1e-300), torch.full_like(w, 1.0 / nb))
    pdf = (1 - eps) * wn + eps / nb
    cdf = torch.cat([torch.zeros((n, 1), dtype=torch.float64), torch.cumsum(pdf, dim=1)], dim=1)
    cdf = cdf / cdf[:, -1:]
    m = K - 2
    u = (torch.arange(m, dtype=torch.float64) + 0.5) / m
    out = torch.empty((n, K), dtype=torch.float64)
    for r in range(n):
        idx = torch.searchsorted(cdf[r], u, right=True) - 1
        idx = idx.clamp(0, nb - 1)
        c0 = cdf[r][idx]
        c1 = cdf[r][idx + 1]
        frac = (u - c0) / (c1 - c0)
        tt = (idx.to(torch.float64) + frac) / nb * float(L[r])
        out[r, 0] = 0.0
        out[r, 1:K - 1] = tt
        out[r, K - 1] = float(L[r])
    return out


def composite_t(rgb, sdf, s):
    """NeuS compositing of one ray with explicit samples (render_torch.py
    155-167 for a single ray): sdf [K], rgb [K-1, 3] at the interval
    midpoints."""
    e = torch.exp(-s * sdf)
    a = torch.clamp((e[1:] - e[:-1]) / (1.0 + e[1:]), 0.0, 10.0)
    T = torch.cat([torch.ones(1, dtype=rgb.dtype), torch.cumprod(1 - a, 0)[:-1]])
    return ((T * a)[:, None] * rgb).sum(0)
