"""Learned warm start: a fully-convolutional 3-D network that maps a lattice's
geometry and a chosen far-field director to an initial Q field.

Every convolution, gradient, loss and optimizer step runs in the library's HIP
kernels (BF16 matrix-core convolutions with FP32 accumulation, qm_ws_* in
include/qmin.h); torch only allocates device buffers and supplies the stream.

Network: four 3x3x3 periodic convolutions with dilations 1, 2, 4, 8 (a
receptive field of 31 sites) and channels 16 -> 32 -> 64 -> 64 -> 16, ReLU
between them; the first five output channels are the correction to the uniform
far-field state 3/2 S0 (d d - I/3), so an untrained network predicts a uniform
start along d.  Inputs (16 channels,
qm_ws_features): site-type one-hot, the anchoring tensor and planar flag and
log W of object sites, and the far-field tensor d d - I/3 of the director the
prediction should adopt far from objects.  Trained on converged lattices
produced by the ensemble minimizer, with the mean squared error to the
converged Q over liquid-crystal sites as the loss.
"""
import ctypes
import math
import time

import numpy as np

from . import QM_FP32, Ensemble, FIRE, Lattice, _check, library
from . import workloads

LAYERS = ((16, 32, 1), (32, 64, 2), (64, 64, 4), (64, 16, 8))   # (cin, cout, dilation)
WCHUNK = 4096    # sites per weight-gradient partial sum
BIAS_BLOCKS = 512   # work-groups (partial sums) of the bias gradient


def _torch():
    import torch
    return torch


def _check_runtime_order():
    from . import check_torch_runtime_order
    check_torch_runtime_order("the warm start is used")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def uniform_q(d, S0):
    """Q of a uniform director d (qTensorFromDirector, 3/2 S0 (d d - I/3))."""
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    return np.array([1.5 * S0 * (d[0] * d[0] - 1 / 3), 1.5 * S0 * d[0] * d[1], 1.5 * S0 * d[0] * d[2],
                     1.5 * S0 * (d[1] * d[1] - 1 / 3), 1.5 * S0 * d[1] * d[2]])


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def far_field_director(q, types):
    """Principal eigenvector of the mean Q over bulk sites (q is [n][5])."""
    bulk = types == 0
    m = q[bulk].mean(axis=0) if bulk.any() else q.mean(axis=0)
    Q = np.array([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], -m[0] - m[3]]])
    w, v = np.linalg.eigh(Q)
    return v[:, int(np.argmax(w))]


class Sample:
    """Device-resident training pair: features, converged Q, site types, and the
    uniform far-field Q the network's output corrects; the geometry's spec, from
    which the energy term of the loss builds a lattice when none is kept."""

    def __init__(self, dims, feats, target, types, offset, lattice=None, spec=None):
        self.dims, self.feats, self.target, self.types = dims, feats, target, types
        self.offset = np.ascontiguousarray(offset, np.float64)
        self.n = dims[0] * dims[1] * dims[2]
        self.n_lc = int((types.cpu().numpy() <= 0).sum()) if hasattr(types, "cpu") else 0
        self.lattice = lattice   # the lattice with this geometry (for the energy term), or None
        self.spec = spec

    def energy_lattice(self):
        """A lattice with this sample's geometry (object sites hold their
        boundary Q; the liquid-crystal sites are overwritten by the prediction)."""
        if self.lattice is not None:
            return self.lattice
        if self.spec is None:
            return None
        S0 = workloads.phase_constants()[3]
        lat = Lattice(*self.spec["dims"], precision=QM_FP32)
        for xyz, plane, wt, W in self.spec.get("walls", []):
            lat.wall(xyz, plane, wt, W, S0)
        for centre, r, btype, W in self.spec["spheres"]:
            lat.sphere(centre, r, btype, W, S0)
        return lat


def lattice_features(lat, director):
    torch = _torch()
    out = torch.empty(lat.n * 16, dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    _check(library().qm_ws_features(lat.h, float(director[0]), float(director[1]), float(director[2]), _ptr(out)))
    return out


def lattice_types(lat):
    torch = _torch()
    out = torch.empty(lat.n, dtype=torch.int8, device="cuda")
    torch.cuda.synchronize()
    _check(library().qm_ws_types(lat.h, _ptr(out)))
    return out


def lattice_q(lat):
    torch = _torch()
    out = torch.empty(lat.n * 5, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _check(library().qm_ws_extract_q(lat.h, _ptr(out)))
    return out


class WarmStartNet:
    def __init__(self, seed=0):
        _check_runtime_order()
        torch = _torch()
        self.lib = library()
        gen = torch.Generator().manual_seed(seed)
        self.w, self.b, self.wbf, self.wbd = [], [], [], []
        self.mw, self.vw, self.mb, self.vb, self.gw, self.gb = [], [], [], [], [], []
        for li, (cin, cout, _) in enumerate(LAYERS):
            std = math.sqrt(2.0 / (27 * cin)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
            w = (torch.randn(27, cout, cin, generator=gen, dtype=torch.float32) * std).cuda()
            self.w.append(w.reshape(-1).contiguous())
            self.b.append(torch.zeros(cout, dtype=torch.float32, device="cuda"))
            for lst, shape, dt in ((self.wbf, 27 * cout * cin, torch.int16), (self.wbd, 27 * cout * cin, torch.int16),
                                   (self.mw, 27 * cout * cin, torch.float32), (self.vw, 27 * cout * cin, torch.float32),
                                   (self.gw, 27 * cout * cin, torch.float32), (self.mb, cout, torch.float32),
                                   (self.vb, cout, torch.float32), (self.gb, cout, torch.float32)):
                lst.append(torch.zeros(shape, dtype=dt, device="cuda"))
        self.step_count = 0
        self.refresh_bf16()

    def refresh_bf16(self):
        for li, (cin, cout, _) in enumerate(LAYERS):
            _check(self.lib.qm_ws_weights_to_bf16(_ptr(self.w[li]), self.w[li].numel(), _ptr(self.wbf[li]),
                                                  _ptr(self.wbd[li]), cout, cin, _stream()))

    # ---------------------------------------------------------------- weights io
    def state(self):
        return {"w%d" % i: self.w[i].cpu().numpy() for i in range(len(LAYERS))} | \
               {"b%d" % i: self.b[i].cpu().numpy() for i in range(len(LAYERS))}

    def save(self, path):
        np.savez(path, layers=np.array(LAYERS), **self.state())

    def load(self, path):
        torch = _torch()
        d = np.load(path)
        if not np.array_equal(d["layers"], np.array(LAYERS)):
            raise ValueError("weights file has another layer layout")
        for i in range(len(LAYERS)):
            self.w[i].copy_(torch.from_numpy(d["w%d" % i]).reshape(-1))
            self.b[i].copy_(torch.from_numpy(d["b%d" % i]))
        self.refresh_bf16()

    # ---------------------------------------------------------------- passes
    def forward(self, feats, dims):
        """Returns the activations [input, y1, y2, y3 (BF16), y4 (FP32 [n][16])]."""
        torch = _torch()
        n = dims[0] * dims[1] * dims[2]
        acts = [feats]
        x = feats
        for li, (cin, cout, dil) in enumerate(LAYERS):
            last = li == len(LAYERS) - 1
            y = torch.empty(n * cout, dtype=torch.float32 if last else torch.int16, device="cuda")
            _check(self.lib.qm_ws_conv_forward(_ptr(x), cin, _ptr(self.wbf[li]), _ptr(self.b[li]), cout, None,
                                               None if last else _ptr(y), _ptr(y) if last else None, dims[0],
                                               dims[1], dims[2], dil, int(li > 0), _stream()))
            acts.append(y)
            x = y
        return acts

    def backward(self, acts, gout, dims):
        torch = _torch()
        n = dims[0] * dims[1] * dims[2]
        chunks = (n + WCHUNK - 1) // WCHUNK
        g = gout
        for li in reversed(range(len(LAYERS))):
            cin, cout, dil = LAYERS[li]
            part = torch.empty(chunks * 27 * cout * cin, dtype=torch.float32, device="cuda")
            _check(self.lib.qm_ws_conv_wgrad(_ptr(g), cout, _ptr(acts[li]), cin, int(li > 0), dims[0], dims[1],
                                             dims[2], dil, _ptr(part), WCHUNK, _ptr(self.gw[li]), _stream()))
            bpart = torch.empty(BIAS_BLOCKS * cout, dtype=torch.float32, device="cuda")
            _check(self.lib.qm_ws_bias_grad(_ptr(g), n, cout, _ptr(bpart), BIAS_BLOCKS, _ptr(self.gb[li]), _stream()))
            if li > 0:
                gin = torch.empty(n * cin, dtype=torch.int16, device="cuda")
                _check(self.lib.qm_ws_conv_forward(_ptr(g), cout, _ptr(self.wbd[li]), None, cin, _ptr(acts[li]),
                                                   _ptr(gin), None, dims[0], dims[1], dims[2], dil, 0, _stream()))
                g = gin

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for li, (cin, cout, _) in enumerate(LAYERS):
            _check(self.lib.qm_ws_adam(_ptr(self.w[li]), _ptr(self.gw[li]), _ptr(self.mw[li]), _ptr(self.vw[li]),
                                       self.w[li].numel(), lr, b1, b2, eps, self.step_count, _ptr(self.wbf[li]),
                                       _ptr(self.wbd[li]), cout, cin, _stream()))
            _check(self.lib.qm_ws_adam(_ptr(self.b[li]), _ptr(self.gb[li]), _ptr(self.mb[li]), _ptr(self.vb[li]),
                                       self.b[li].numel(), lr, b1, b2, eps, self.step_count, None, None, cout, 0,
                                       _stream()))

    def loss_and_grad(self, y, sample):
        torch = _torch()
        blocks = 512
        g = torch.empty(sample.n * 16, dtype=torch.int16, device="cuda")
        part = torch.empty(blocks, dtype=torch.float64, device="cuda")
        scale = 1.0 / max(sample.n_lc * 5, 1)
        _check(self.lib.qm_ws_mse_grad(_ptr(y), 16, _ptr(sample.target), _ptr(sample.types), sample.n, scale,
                                       _dp(sample.offset), _ptr(g), _ptr(part), blocks, _stream()))
        return g, part

    def energy_grad(self, y, g, sample, force, weight):
        """Adds the gradient of the energy term of the loss, weight x E(prediction)
        / (5 n_lc), to g: the prediction goes into a lattice of the sample's
        geometry, whose force F (minus the energy gradient, shaped by the metric,
        from the library's force pass) adds -weight F / (5 n_lc).  Returns the
        term's value (energy per liquid-crystal component, times weight)."""
        lat = sample.energy_lattice()
        if lat is None:
            return 0.0
        _torch().cuda.synchronize()
        _check(self.lib.qm_ws_apply(lat.h, _ptr(y), 16, _dp(sample.offset)))
        scale = weight / max(sample.n_lc * 5, 1)
        _check(self.lib.qm_ws_energy_grad(force.h, lat.h, scale, _ptr(g), 16, _stream()))
        e, _ = force.compute_energy(lat)
        return weight * e / max(sample.n_lc * 5, 1)

    def train_step(self, sample, lr=1e-3, force=None, energy_weight=0.0):
        acts = self.forward(sample.feats, sample.dims)
        g, part = self.loss_and_grad(acts[-1], sample)
        self.last_energy_term = 0.0
        if energy_weight > 0 and force is not None:
            self.last_energy_term = self.energy_grad(acts[-1], g, sample, force, energy_weight)
        self.backward(acts, g, sample.dims)
        self.adam_step(lr)
        return part

    def train(self, samples, epochs=4, lr=1e-3, seed=0, log=None, force=None, energy_weight=0.0):
        """Adam over the samples; the loss is the mean squared error to the
        converged Q plus, with energy_weight > 0 and samples that keep their
        lattice, energy_weight times the Landau-de Gennes energy of the
        prediction per liquid-crystal component (force: the force computer)."""
        torch = _torch()
        rng = np.random.default_rng(seed)
        history = []
        for ep in range(epochs):
            order = rng.permutation(len(samples))
            parts, eterm = [], 0.0
            for k in order:
                parts.append(self.train_step(samples[k], lr, force, energy_weight))
                eterm += self.last_energy_term
            torch.cuda.synchronize()
            mse = sum(float(p.sum().item()) / max(samples[k].n_lc * 5, 1) for p, k in zip(parts, order)) / len(order)
            history.append(mse)
            if log:
                log("epoch %d: mean squared error %.4e, energy term %.6e" % (ep + 1, mse, eterm / len(order)))
        return history

    def predict(self, lat, director, S0=None):
        """Predicted Q written into the lattice's liquid-crystal sites."""
        torch = _torch()
        if S0 is None:
            S0 = workloads.phase_constants()[3]
        feats = lattice_features(lat, director)
        acts = self.forward(feats, lat.shape)
        torch.cuda.synchronize()
        _check(self.lib.qm_ws_apply(lat.h, _ptr(acts[-1]), 16, _dp(uniform_q(director, S0))))
        return acts[-1]


# ---------------------------------------------------------------- geometries
def diagonal_director():
    """Far-field director of predictions on periodic lattices: a body diagonal
    of the cubic lattice.  The discrete operators are not isotropic, and around
    objects the texture's energy depends on how the far field sits against the
    lattice axes; minimized states on periodic lattices end with far fields near
    the body diagonals, and the far field turns towards them only slowly (a soft
    mode: a start along z on a 128^3 lattice with five spheres took 104k FIRE
    iterations, one along the far field it ended with 15.6k).  Predicting along a
    diagonal skips most of that turn."""
    return tuple(float(v) for v in np.ones(3) / math.sqrt(3.0))


def heldout_geometry(L, seed):
    """Spheres of a held-out periodic L^3 geometry: 2-6 spheres at seeded
    positions, radius uniform in [L/16, L/8], homeotropic or degenerate planar,
    W log-uniform in [0.5, 20]: [(centre, radius, btype, W)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(int(rng.integers(2, 7))):
        c = tuple(float(v) for v in rng.uniform(0, L, 3))
        out.append((c, float(rng.uniform(L / 16, L / 8)), int(rng.integers(0, 2)),
                    float(math.exp(rng.uniform(math.log(0.5), math.log(20.0))))))
    return out


def diagonal_start_specs(count, size=64, seed=8192):
    """The ensemble workload's geometries at size^3 (as random_start_specs)
    started from a uniform director along diagonal_director(): targets in the
    far-field orientation predictions on periodic lattices use."""
    specs = workloads.ensemble_specs(count, seed, edges=(size,))
    for sp in specs:
        sp["director"] = diagonal_director()
    return specs


def random_start_specs(count, size=64, seed=4096):
    """Training geometries as the ensemble workload's at size^3 (periodic, 0-3
    spheres of radius 4-16, homeotropic or degenerate planar, W log-uniform in
    [0.5, 20]) with seeded random-director starts: the targets are the states
    the ensemble minimizer reaches from them."""
    return workloads.ensemble_specs(count, seed, edges=(size,))


# ---------------------------------------------------------------- training data
def training_specs(count, size=64, seed=77):
    """Geometries of the training set: size^3 periodic lattices, half of them
    with z walls (either anchoring), 0-3 spheres as the ensemble workload, and a
    seeded initial director the minimization starts from uniformly (random on
    the sphere; in the wall plane for planar walls, along z for homeotropic
    ones), so that each converged target is a function of the geometry and the
    far-field director rather than of a random start."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        d = rng.normal(size=3)
        spec = {"dims": (size, size, size), "seed": int(rng.integers(1, 2 ** 31 - 1)), "spheres": [], "walls": []}
        if rng.random() < 0.5:
            wt = int(rng.integers(0, 2))
            W = float(math.exp(rng.uniform(math.log(0.5), math.log(20.0))))
            spec["walls"] = [(2, 0, wt, W), (2, size - 1, wt, W)]
            d = np.array([d[0], d[1], 0.0]) if wt == 1 else np.array([0.0, 0.0, 1.0])
        spec["director"] = tuple(float(v) for v in d / np.linalg.norm(d))
        for _ in range(int(rng.integers(0, 4))):
            r = float(min(rng.uniform(4.0, 16.0), size / 3.0))
            centre = tuple(float(rng.uniform(0, size)) for _ in range(3))
            spec["spheres"].append((centre, r, int(rng.integers(0, 2)), float(math.exp(
                rng.uniform(math.log(0.5), math.log(20.0))))))
        specs.append(spec)
    return specs


def build_lattice(qmin, spec, precision, S0):
    lat = qmin.Lattice(*spec["dims"], precision=precision)
    if "director" in spec:
        lat.uniform_init(spec["director"], S0)
    else:
        lat.random_init(spec["seed"], S0)
    for xyz, plane, wt, W in spec.get("walls", []):
        lat.wall(xyz, plane, wt, W, S0)
    for centre, r, btype, W in spec["spheres"]:
        lat.sphere(centre, r, btype, W, S0)
    return lat


def make_dataset(qmin, specs, force, fc=1e-7, max_iter=4000, batch=128, log=None, keep_lattices=False):
    """Minimize the geometries with the ensemble minimizer (FP32) and keep
    device-resident training samples (with their lattices, for the energy term
    of the loss, when keep_lattices)."""
    A, B, C, S0 = workloads.phase_constants()
    samples = []
    for start in range(0, len(specs), batch):
        part = specs[start:start + batch]
        lats = [build_lattice(qmin, s, QM_FP32, S0) for s in part]
        t0 = time.perf_counter()
        its, fms = Ensemble(lats).fire_minimize(FIRE(dt=5e-4, dt_max=2e-2, fc=fc, max_iter=max_iter), force, 32)
        for lat, spec in zip(lats, part):
            q = lat.q()
            types = lat.types()
            d = far_field_director(q, types)
            s = Sample(lat.shape, lattice_features(lat, d), lattice_q(lat), lattice_types(lat), uniform_q(d, S0),
                       lat if keep_lattices else None, spec)
            samples.append(s)
        if log:
            log("minimized %d training lattices in %.1f s (iterations mean %.0f, converged %d)"
                % (len(part), time.perf_counter() - t0, float(np.mean(its)), int((fms <= fc).sum())))
    return samples
