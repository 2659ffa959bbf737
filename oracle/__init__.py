"""CPU oracle for the QSDVR render hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``qsdvr_amd``)
imports this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Modules
-------
raygen   bit-exact float32 restatement of kernel/camera.cu (RayCut kernel,
         the host exclusive scan, GenerateRayPoints).
render   restatement of models/render_torch.py (quadric grid, SDF, normal,
         trilinear appearance interpolation, degree-2 SH shading, NeuS
         logistic compositing), dtype-generic (float32 / float64) torch-CPU
         code whose autograd gives the reference gradient.
mlp      float64 restatement of the proposal MLP (positional encoding,
         39->64->64->64->1 ReLU network) and of the inverse-CDF sampler.

Pinning (what checks each function independently of its own formula) is
listed in each module's docstring; every function is pinned by closed forms,
invariants, finite differences or brute-force loops in ``tests/``.
"""
