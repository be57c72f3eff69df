"""Learned site-PSF surrogate with audit and fallback (no counterpart in the reference).

A convolutional encoder-decoder maps a site's downsampled refractive-index
patch (8 depth bins x 64 x 64, plus a focal-depth channel) to its local
two-photon PSF (9 planes x 33 x 33 around focus).  Every layer runs on the
hand-written kernels of libmaomi_amd.so: im2col + BF16 MFMA GEMM with FP32
accumulation for the forward and both backward products, bias + leaky ReLU,
nearest 2x upsampling, a relative squared-error loss and Adam on FP32 master
weights with BF16 copies.  The training data are site PSFs produced by the
propagator itself.  At inference a seeded random fraction of sites is re-run
with the FP32 propagator; an audited site whose prediction is off by more than
the tolerance is replaced by the propagator, and when too many audited sites
fail the whole run falls back to the propagator.
"""

import numpy as np
import torch

from . import _native as nat


def _st(dev):
    return nat.stream_handle(dev)


def gemm(A, B, C, M, N, K, sam, sak, sbk, sbn, ldc, alpha=1.0, beta=0.0):
    nat.check(nat.lib().maomi_gemm_bf16(nat.ptr(A), nat.ptr(B), nat.ptr(C), M, N, K, sam, sak,
                                        sbk, sbn, ldc, float(alpha), float(beta),
                                        _st(C.device)), "gemm bf16")


def to_bf16(x, out=None):
    out = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device) if out is None else out
    x = x.contiguous()  # kept alive until the launch is enqueued
    nat.check(nat.lib().maomi_to_bf16(nat.ptr(x), nat.ptr(out), x.numel(),
                                      _st(x.device)), "to bf16")
    return out


class Conv:
    """k x k convolution, stride s, padding p, activation leaky ReLU(slope) (slope 1: linear)."""

    def __init__(self, cin, cout, k, s, p, slope, gen, device):
        self.cin, self.cout, self.k, self.s, self.p, self.slope = cin, cout, k, s, p, slope
        fan = cin * k * k
        std = np.sqrt(2.0 / fan) if slope < 1 else np.sqrt(1.0 / fan)
        w = torch.randn((cout, fan), generator=gen, dtype=torch.float32) * std
        self.w = w.to(device)
        self.b = torch.zeros(cout, dtype=torch.float32, device=device)
        self.wb = to_bf16(self.w)
        self.gw, self.gb = torch.zeros_like(self.w), torch.zeros_like(self.b)
        self.state = [torch.zeros_like(self.w), torch.zeros_like(self.w),
                      torch.zeros_like(self.b), torch.zeros_like(self.b)]

    def params(self):
        return self.w.numel() + self.b.numel()

    def out_hw(self, H, W):
        return (H + 2 * self.p - self.k) // self.s + 1, (W + 2 * self.p - self.k) // self.s + 1

    def forward(self, x, train):
        C, B, H, W = x.shape
        Ho, Wo = self.out_hw(H, W)
        n = B * Ho * Wo
        K = self.cin * self.k * self.k
        col = torch.empty((K, n), dtype=torch.bfloat16, device=x.device)
        nat.check(nat.lib().maomi_im2col_bf16(nat.ptr(x), nat.ptr(col), C, B, H, W, self.k, self.s,
                                              self.p, _st(x.device)), "im2col")
        y = torch.empty((self.cout, B, Ho, Wo), dtype=torch.float32, device=x.device)
        gemm(self.wb, col, y, self.cout, n, K, K, 1, n, 1, n)
        nat.check(nat.lib().maomi_bias_act(nat.ptr(y), nat.ptr(self.b), self.cout, n, self.slope,
                                           _st(x.device)), "bias act")
        if train:
            self.cache = (x.shape, col, y)
        return y

    def backward(self, dy):
        (C, B, H, W), col, y = self.cache
        Ho, Wo = y.shape[2], y.shape[3]
        n = B * Ho * Wo
        K = self.cin * self.k * self.k
        nat.check(nat.lib().maomi_act_bwd(nat.ptr(dy), nat.ptr(y), self.cout, n, self.slope,
                                          nat.ptr(self.gb), _st(dy.device)), "act bwd")
        dyb = to_bf16(dy)
        # dW[co][kk] = sum_n dY[co][n] col[kk][n]
        gemm(dyb, col, self.gw, self.cout, K, n, n, 1, 1, n, K)
        # dcol[kk][n] = sum_co W[co][kk] dY[co][n]
        dcol = torch.empty((K, n), dtype=torch.float32, device=dy.device)
        gemm(self.wb, dyb, dcol, K, n, self.cout, 1, K, n, 1, n)
        dx = torch.empty((C, B, H, W), dtype=torch.float32, device=dy.device)
        nat.check(nat.lib().maomi_col2im(nat.ptr(dcol), nat.ptr(dx), C, B, H, W, self.k, self.s,
                                         self.p, _st(dy.device)), "col2im")
        self.cache = None
        return dx

    def step(self, lr, t, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
        lib = nat.lib()
        st = _st(self.w.device)
        m, v, mb, vb = self.state
        nat.check(lib.maomi_adam(nat.ptr(self.w), nat.ptr(self.wb), nat.ptr(self.gw), nat.ptr(m),
                                 nat.ptr(v), self.w.numel(), lr, b1, b2, eps, t, wd, st), "adam w")
        nat.check(lib.maomi_adam(nat.ptr(self.b), None, nat.ptr(self.gb), nat.ptr(mb), nat.ptr(vb),
                                 self.b.numel(), lr, b1, b2, eps, t, 0.0, st), "adam b")


class Template:
    """Learned output planes T (C, H, W) added to every sample's output; the network then
    predicts the site-specific part on top of T.  Set to the training-mean PSF before training."""

    def __init__(self, C, H, W, device):
        self.C, self.H, self.W = C, H, W
        self.w = torch.zeros((C, H * W), dtype=torch.float32, device=device)
        self.gw = torch.zeros_like(self.w)
        self.state = [torch.zeros_like(self.w), torch.zeros_like(self.w)]

    def params(self):
        return self.w.numel()

    def set_mean(self, t):
        """T = mean over samples of channel-major targets t (C, P, H, W)."""
        C, P, H, W = t.shape
        t = t.contiguous()
        nat.check(nat.lib().maomi_plane_bias_grad(nat.ptr(t), nat.ptr(self.w), C, P,
                                                  H * W, 1.0 / P, _st(t.device)), "template mean")

    def forward(self, x, train):
        C, B, H, W = x.shape
        if (C, H, W) != (self.C, self.H, self.W):
            raise ValueError(f"template planes {(self.C, self.H, self.W)} vs output {(C, H, W)}")
        y = torch.empty_like(x)
        nat.check(nat.lib().maomi_plane_bias(nat.ptr(x), nat.ptr(self.w), nat.ptr(y), C, B, H * W,
                                             _st(x.device)), "template")
        return y

    def backward(self, dy):
        C, B, H, W = dy.shape
        nat.check(nat.lib().maomi_plane_bias_grad(nat.ptr(dy), nat.ptr(self.gw), C, B, H * W, 1.0,
                                                  _st(dy.device)), "template grad")
        return dy

    def step(self, lr, t, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
        m, v = self.state
        nat.check(nat.lib().maomi_adam(nat.ptr(self.w), None, nat.ptr(self.gw), nat.ptr(m),
                                       nat.ptr(v), self.w.numel(), lr, b1, b2, eps, t, 0.0,
                                       _st(self.w.device)), "adam template")


class Up2:
    def forward(self, x, train):
        C, B, H, W = x.shape
        y = torch.empty((C, B, 2 * H, 2 * W), dtype=torch.float32, device=x.device)
        nat.check(nat.lib().maomi_upsample2(nat.ptr(x), nat.ptr(y), C * B, H, W, 0, _st(x.device)),
                  "upsample")
        self.shape = x.shape
        return y

    def backward(self, dy):
        C, B, H, W = self.shape
        dx = torch.empty((C, B, H, W), dtype=torch.float32, device=dy.device)
        nat.check(nat.lib().maomi_upsample2(nat.ptr(dy), nat.ptr(dx), C * B, H, W, 1,
                                            _st(dy.device)), "upsample bwd")
        return dx

    def params(self):
        return 0

    def step(self, *a, **k):
        pass


def default_layers(cin=11, widths=(32, 64, 128, 384, 384, 96, 48), out_ch=9, side=33,
                   slope=0.1):
    """Encoder 64 -> 32 -> 16 -> 8 (two 8x8 convs), decoder 8 -> 16 -> 32, 2x2 conv to 33x33,
    plus learned output planes.  Input: 8 index bins, depth, and two coordinate channels (the
    convolutions alone cannot place the focal spot at the window centre)."""
    w = widths
    return [("conv", cin, w[0], 3, 2, 1, slope), ("conv", w[0], w[1], 3, 2, 1, slope),
            ("conv", w[1], w[2], 3, 2, 1, slope), ("conv", w[2], w[3], 3, 1, 1, slope),
            ("conv", w[3], w[4], 3, 1, 1, slope), ("up",), ("conv", w[4], w[5], 3, 1, 1, slope),
            ("up",), ("conv", w[5], w[6], 3, 1, 1, slope), ("conv", w[6], out_ch, 2, 1, 1, 1.0),
            ("template", out_ch, side, side)]


def _make_layer(l, gen, device):
    if l[0] == "up":
        return Up2()
    if l[0] == "template":
        return Template(*l[1:], device=device)
    return Conv(*l[1:], gen=gen, device=device)


class SurrogateNet:
    def __init__(self, layers=None, seed=0, device="cuda"):
        self.device = torch.device(device)
        gen = torch.Generator().manual_seed(seed)
        self.spec = layers or default_layers()
        self.layers = [_make_layer(l, gen, self.device) for l in self.spec]
        if isinstance(self.layers[-1], Template):
            # the convolutional part starts near zero: the net starts at the template
            last = self.layers[-2]
            last.w.mul_(0.1)
            last.wb = to_bf16(last.w)
        self.t = 0

    @property
    def template(self):
        return self.layers[-1] if isinstance(self.layers[-1], Template) else None

    def num_params(self):
        return sum(l.params() for l in self.layers)

    def forward(self, x, train=False):
        for l in self.layers:
            x = l.forward(x, train)
        return x

    def backward(self, dy):
        for l in reversed(self.layers):
            dy = l.backward(dy)
        return dy

    def loss(self, y, t, grad=True, kind="energy"):
        """Per-sample loss and its gradient.  "energy": |y - t|^2 over the batch's mean target
        energy (training: low-energy sites do not dominate); "relative": |y - t|^2 / |t|^2."""
        C, B, H, W = y.shape
        g = torch.empty_like(y) if grad else None
        per = torch.empty(B, dtype=torch.float32, device=y.device)
        if kind == "energy":
            work = torch.empty(2 * B, dtype=torch.float64, device=y.device)
            nat.check(nat.lib().maomi_energy_sq_loss(nat.ptr(y), nat.ptr(t), nat.ptr(g),
                                                     nat.ptr(per), nat.ptr(work), C, B, H * W,
                                                     _st(y.device)), "energy loss")
        else:
            nat.check(nat.lib().maomi_rel_sq_loss(nat.ptr(y), nat.ptr(t), nat.ptr(g), nat.ptr(per),
                                                  C, B, H * W, _st(y.device)), "loss")
        return per, g

    def train_step(self, x, t, lr, wd=0.0):
        y = self.forward(x, train=True)
        per, g = self.loss(y, t)
        self.backward(g)
        self.t += 1
        for l in self.layers:
            l.step(lr, self.t, wd)
        return per

    def forward_flops(self, H, W):
        """Multiply-add flops (2 per MAC) of one sample's forward pass on an H x W input."""
        f = 0
        for l in self.layers:
            if isinstance(l, Conv):
                H, W = l.out_hw(H, W)
                f += 2 * l.cout * l.cin * l.k * l.k * H * W
            elif isinstance(l, Up2):
                H, W = 2 * H, 2 * W
        return f

    def predict(self, x, batch=4096):
        """Inference: channel-last BF16 activations through the implicit-GEMM convolution
        (maomi_conv_bf16_fwd: no im2col buffer, bias and leaky ReLU on the accumulators, the
        decoder's upsampling folded into the next convolution's reads).  x: (C, P, H, W)."""
        outs = []
        for s in range(0, x.shape[1], batch):
            outs.append(self._infer(x[:, s:s + batch].contiguous()))
        return torch.cat(outs, 1) if len(outs) > 1 else outs[0]

    def _channel_last_weights(self):
        """Per convolution: BF16 weights [co][ky][kx][c] with c padded to a multiple of 8 (the
        same BF16 values as the training path's copies), rebuilt when the weights change."""
        if getattr(self, "_cl_t", None) != self.t:
            self._cl = []
            for l in self.layers:
                if isinstance(l, Conv):
                    cp = -(-l.cin // 8) * 8
                    w = l.wb.view(l.cout, l.cin, l.k, l.k).permute(0, 2, 3, 1)
                    wp = torch.zeros((l.cout, l.k, l.k, cp), dtype=torch.bfloat16,
                                     device=self.device)
                    wp[..., :l.cin] = w
                    self._cl.append(wp.reshape(l.cout, -1).contiguous())
            self._cl_t = self.t
        return self._cl

    def _infer(self, x):
        lib = nat.lib()
        st = _st(x.device)
        C, B, H, W = x.shape
        cp = -(-C // 8) * 8
        act = torch.empty((B, H, W, cp), dtype=torch.bfloat16, device=x.device)
        nat.check(lib.maomi_to_channel_last_bf16(nat.ptr(x), C, B, H * W, cp, nat.ptr(act), st),
                  "channel-last input")
        weights = iter(self._channel_last_weights())
        convs = [i for i, l in enumerate(self.layers) if isinstance(l, Conv)]
        up, out = 0, None
        for i, l in enumerate(self.layers):
            if isinstance(l, Up2):
                up = 1
                continue
            if isinstance(l, Template):
                out = l.forward(out, False)
                continue
            last = i == convs[-1]
            Hi, Wi = H << up, W << up
            Ho, Wo = (Hi + 2 * l.p - l.k) // l.s + 1, (Wi + 2 * l.p - l.k) // l.s + 1
            if last:
                y = torch.empty((l.cout, B, Ho, Wo), dtype=torch.float32, device=x.device)
            else:
                if l.cout % 8:
                    raise ValueError("hidden convolutions need a multiple of 8 channels")
                y = torch.empty((B, Ho, Wo, l.cout), dtype=torch.bfloat16, device=x.device)
            nat.check(lib.maomi_conv_bf16_fwd(nat.ptr(act), B, H, W, act.shape[-1], up,
                                              nat.ptr(next(weights)), nat.ptr(l.b), l.cout, l.k,
                                              l.s, l.p, float(l.slope), nat.ptr(y), int(not last),
                                              st), "implicit conv")
            act, H, W, up = y, Ho, Wo, 0
            if last:
                out = y
        return out


def to_cbhw(a):
    """(B, C, H, W) -> channel-major (C, B, H, W) contiguous."""
    return a.permute(1, 0, 2, 3).contiguous()


def train(net, x, t, *, epochs=10, batch=64, lr=1e-3, wd=0.0, seed=0, warmup=0.05, log=None):
    """Adam with linear warm-up over the first `warmup` fraction of steps, then cosine decay
    per epoch.  x, t: channel-major tensors (C, P, H, W) on the device."""
    P = x.shape[1]
    if P < batch:
        raise ValueError(f"{P} training samples for batch {batch}")
    rng = np.random.default_rng(seed)
    if net.template is not None and net.t == 0:
        net.template.set_mean(t)
    steps = epochs * (P // batch)
    warm = max(1, int(warmup * steps))
    done = 0
    hist = []
    for ep in range(epochs):
        order = torch.as_tensor(rng.permutation(P), device=x.device)
        tot, cnt = 0.0, 0
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / max(1, epochs)))
        for s in range(0, P - batch + 1, batch):
            idx = order[s:s + batch]
            done += 1
            per = net.train_step(x[:, idx].contiguous(), t[:, idx].contiguous(),
                                 lr_ep * min(1.0, done / warm), wd)
            tot += float(per.sum())
            cnt += batch
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: mean rel sq error {hist[-1]:.4e} (lr {lr_ep:.2e})")
    return hist


def rel_l2_per_site(pred, true):
    """Per-site relative L2 |pred - true| / |true| of channel-major tensors (C, P, H, W)."""
    C, P, H, W = true.shape
    per = torch.empty(P, dtype=torch.float32, device=true.device)
    # both operands bound to names: a temporary freed before the launch can hand its block to
    # the next temporary
    pred, true = pred.contiguous(), true.contiguous()
    nat.check(nat.lib().maomi_rel_sq_loss(nat.ptr(pred), nat.ptr(true), None, nat.ptr(per), C, P,
                                          H * W, _st(true.device)), "rel l2")
    return per.sqrt_()


# ---------------------------------------------------------------------- data and stage
PSF_PLANES, PSF_SIDE, PATCH_POOL, PATCH_BINS = 9, 33, 4, 8


def site_features(volume, cfg, offsets, depth_um, *, pool=PATCH_POOL, zbins=PATCH_BINS):
    """Network input (zbins + 3, P, N/pool, N/pool): pooled (n - n0)/std over each site window,
    the depth scalar, then x and y coordinates in [-1, 1]."""
    d = volume.device
    offs = np.asarray(offsets, np.int64)
    P = len(offs)
    wx = torch.as_tensor(offs[:, 0].astype(np.int32), device=d)
    wy = torch.as_tensor(offs[:, 1].astype(np.int32), device=d)
    dep = torch.full((P,), float(depth_um) / 100.0, dtype=torch.float32, device=d)
    S = cfg.n // pool
    out = torch.empty((zbins + 3, P, S, S), dtype=torch.float32, device=d)
    ramp = torch.linspace(-1.0, 1.0, S, dtype=torch.float32, device=d)
    out[zbins + 1].copy_(ramp[None, :, None].expand(P, S, S))
    out[zbins + 2].copy_(ramp[None, None, :].expand(P, S, S))
    X, Y, Zv = volume.shape
    zsum = torch.empty((zbins, X, Y), dtype=torch.float64, device=d)
    nat.check(nat.lib().maomi_patch_features(nat.ptr(volume), X, Y, Zv, nat.ptr(wx), nat.ptr(wy),
                                             nat.ptr(dep), P, cfg.n, pool, zbins, float(cfg.n0),
                                             1.0 / cfg.dn_std, nat.ptr(zsum), nat.ptr(out),
                                             _st(d)), "patch features")
    return out


def focal_window(cfg, planes=PSF_PLANES, side=PSF_SIDE):
    """(z0, nz, c0, w): planes centred on the focal plane, crop centred on the site."""
    zf = int(round(cfg.focus / cfg.dz_um))
    return zf - planes // 2, planes, cfg.n // 2 - side // 2, side


def site_psfs(stage, subset=None):
    """Propagator PSFs (nz, P, w, w) of all sites of a planned PlainStage (or a subset)."""
    z0, nz, c0, w = focal_window(stage.cfg)
    P = len(stage.offsets) if subset is None else len(subset)
    out = torch.empty((nz, P, w, w), dtype=torch.float32, device=stage.device)
    stage.propagate(subset=subset, psf=(out, z0, nz, c0, w))
    return out


class SurrogateStage:
    """Surrogate inference with a seeded audit inside the propagator's launch.

    The network predicts the audited sites (a seeded fraction) first.  One persistent
    focal-plane launch then propagates the sites with the audited ones first in its order; as
    each audited site is done its prediction is compared on the device (maomi_sites_audit).
    Once every audited site is done with at most `max_fail_frac` of them above `tol`, the
    launch's blocks stop taking unaudited sites and the network predicts those that were left.
    With more misses the launch simply propagates every site: the fallback costs the audit's
    inference and no second launch.  Every site the launch propagated keeps its propagated PSF
    (audited sites included)."""

    def __init__(self, net, scale, stage, *, audit_frac=0.05, tol=0.01, max_fail_frac=0.1,
                 seed=0, schedule="static"):
        self.net, self.scale, self.stage = net, float(scale), stage
        self.schedule = schedule  # of the focal-plane launch: "static" or "queue"
        self.audit_frac, self.tol, self.max_fail_frac = audit_frac, tol, max_fail_frac
        self.rng = np.random.default_rng(seed)
        self.timing, self.report = {}, {}

    def run(self):
        import time
        st, cfg = self.stage, self.stage.cfg
        P = len(st.offsets)
        dev = st.device
        t0 = time.perf_counter()
        k = max(1, int(np.ceil(self.audit_frac * P)))
        audit = np.sort(self.rng.choice(P, k, replace=False))
        rest = np.setdiff1d(np.arange(P), audit)
        x_audit = site_features(st.volume, cfg, st.offsets[audit], cfg.focus)
        pred = (self.net.predict(x_audit) * self.scale).contiguous()
        slot = np.full(P, -1, np.int32)
        slot[audit] = np.arange(k, dtype=np.int32)
        au = {"pred": pred, "slot": torch.as_tensor(slot, device=dev),
              "state": torch.zeros(2, dtype=torch.int32, device=dev),
              "rel": torch.empty(k, dtype=torch.float32, device=dev),
              "done": torch.zeros(P, dtype=torch.int32, device=dev),
              "max_fail": int(np.floor(self.max_fail_frac * k)), "tol": self.tol}
        z0, nz, c0, w = focal_window(cfg)
        out = torch.empty((nz, P, w, w), dtype=torch.float32, device=dev)
        st.propagate(psf=(out, z0, nz, c0, w), order=np.concatenate([audit, rest]), audit=au,
                     schedule=self.schedule)
        state = au["state"].cpu().numpy()
        done = au["done"].cpu().numpy().astype(bool)
        rel = au["rel"].cpu().numpy()
        t1 = time.perf_counter()
        fails = int(state[1])
        if int(state[0]) != k:
            raise RuntimeError(f"audit incomplete: {int(state[0])} of {k} audited sites done")
        mode = "surrogate" if fails <= au["max_fail"] else "fallback_all"
        left = np.flatnonzero(~done)
        if mode == "fallback_all" and len(left):
            raise RuntimeError("a failed audit must leave every site propagated")
        if len(left):
            lidx = torch.as_tensor(left, device=dev)
            x = site_features(st.volume, cfg, st.offsets[left], cfg.focus)
            out[:, lidx] = self.net.predict(x) * self.scale
        torch.cuda.synchronize(dev)
        t2 = time.perf_counter()
        self.timing = {"audit_and_propagation": t1 - t0, "predicted_sites": t2 - t1,
                       "total": t2 - t0}
        self.report = {"mode": mode, "audited": int(k), "audit_fail": fails,
                       "propagated_sites": int(done.sum()), "predicted_sites": int(len(left)),
                       "audit_rel_l2_median": float(np.median(rel)),
                       "audit_rel_l2_max": float(rel.max()),
                       "audit_rel_l2_mean": float(rel.mean())}
        self.psfs = out
        return out
