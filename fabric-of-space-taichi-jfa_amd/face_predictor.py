"""Learned face predictor: host side over the fos_pred_* / fos_mlp_* kernels.

No reference counterpart.  On frames where the JFA is decimated it predicts,
for every hash-grid candidate pair (i, j) within MULT*(ri+rj) (a flat list,
each unordered pair once), whether i and j share an accepted power-diagram
face, and from that the FSC of every
particle and the face rows the pressure step reads.  The plain JFA produces
the training labels (its neighbour lists) and stays the judge.

Parameters are one flat fp32 vector (W1[16][64] b1[64] W2[64][64] b2[64]
w3[64] b3; rows 13-15 of W1 multiply the zero feature slots).  Feature 12 is
the hysteresis the JFA's finalize applies: whether the pair was accepted in
the previous frame (``prev_up``: the JFA's neighbour lists after a JFA frame,
the predicted face rows after a predicted one).  Inference
packs W1 and W2 to BF16 in MFMA lane order (``pack``).
"""
import ctypes
import math

import numpy as np
import torch

import fos_lib as fl

H = 64
NFEAT = 13
SLOTS = 16
OFF_W1, OFF_B1 = 0, 16 * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1
MULT = 1.6
CAP = 96
CLIP = 1.5      # the face polygon starts as a square of half-size CLIP*(ri+rj)


def param_count():
    c = ctypes.c_int()
    fl.call("fos_mlp_param_count", ctypes.byref(c))
    return c.value


def init_params(seed=0):
    """He-initialised fp32 parameters (numpy)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    w1 = np.zeros((16, H), np.float32)
    w1[:NFEAT] = rng.normal(0, math.sqrt(2.0 / NFEAT), (NFEAT, H))
    p[OFF_W1:OFF_B1] = w1.ravel()
    p[OFF_W2:OFF_B2] = rng.normal(0, math.sqrt(2.0 / H), (H, H)).ravel()
    p[OFF_W3:OFF_B3] = rng.normal(0, math.sqrt(1.0 / H), H)
    return p


def bf16_bits(x):
    """fp32 -> bf16 bit patterns, round to nearest even (as the kernels do)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_round(x):
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float32)
    return (p[OFF_W1:OFF_B1].reshape(16, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], float(p[OFF_B3]))


def pack(p, device="cuda"):
    """Device tensors for fos_pred_mlp: W1/W2 as BF16 in MFMA A-operand lane
    order (lane l of block (mb, kb) holds W[16kb + 4(l/16) + v][16mb + l%16],
    v = 0..3), biases and the last layer in fp32."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    lane = np.arange(64)
    col, grp = lane % 16, lane // 16
    w1p = np.zeros((4, 64, 4), np.float32)
    w2p = np.zeros((4, 4, 64, 4), np.float32)
    for mb in range(4):
        for v in range(4):
            w1p[mb, :, v] = w1[4 * grp + v, 16 * mb + col]
            for kb in range(4):
                w2p[mb, kb, :, v] = w2[16 * kb + 4 * grp + v, 16 * mb + col]
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)
    return {
        "w1p": t(bf16_bits(w1p).view(np.int16), torch.int16),
        "w2p": t(bf16_bits(w2p).view(np.int16), torch.int16),
        "b1": t(b1, torch.float32), "b2": t(b2, torch.float32), "w3": t(w3, torch.float32),
        "b3": float(b3),
    }


def mlp_reference(feat, p):
    """numpy reference of the inference kernel: BF16 inputs, weights and
    hidden activations, fp32 accumulation."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = bf16_round(np.asarray(feat, np.float32).reshape(-1, SLOTS))
    h1 = np.maximum(x @ bf16_round(w1) + b1, 0).astype(np.float32)
    h2 = np.maximum(bf16_round(h1) @ bf16_round(w2) + b2, 0).astype(np.float32)
    return (h2 @ w3 + b3).astype(np.float32)


def hash_res(L, r_max, mult=MULT):
    """Cells of at least mult * 2 * r_max (the 27-cell stencil reaches every
    candidate)."""
    return max(3, int(math.floor(L / (mult * 2.0 * r_max))))


class FacePredictor:
    """Scores candidate pairs of one foam; owns its hash grid and buffers."""

    def __init__(self, params, n_max, L, r_max, device="cuda", cap=CAP, mult=MULT):
        fl.require_gpu()
        self.device = torch.device(device)
        self.L, self.cap, self.mult = float(L), int(cap), float(mult)
        self.G = hash_res(L, r_max, mult)
        self.n_max = int(n_max)
        self.ctx = fl.Context(self.n_max, 1, self.G ** 3)
        d, i32, f32 = self.device, torch.int32, torch.float32
        self.pos = torch.empty((self.n_max, 3), dtype=f32, device=d)
        self.cc = torch.empty(self.G ** 3, dtype=i32, device=d)
        self.cs = torch.empty(self.G ** 3, dtype=i32, device=d)
        self.ci = torch.empty(self.n_max, dtype=i32, device=d)
        self.cand = torch.empty((self.n_max, self.cap), dtype=i32, device=d)
        self.ncand = torch.empty(self.n_max, dtype=i32, device=d)
        self.nup = torch.empty(self.n_max, dtype=i32, device=d)
        self.trunc = torch.empty(self.n_max, dtype=i32, device=d)
        self.rel = torch.empty((self.n_max, self.cap, 4), dtype=f32, device=d)
        self.pair_off = torch.empty(self.n_max + 1, dtype=i32, device=d)
        self.pair_i = torch.empty(self.n_max * self.cap, dtype=i32, device=d)
        self.feat = torch.empty((self.n_max * self.cap, SLOTS), dtype=f32, device=d)
        # previous frame's accepted partners j > i per row (-1: none yet);
        # warm: prev_up holds a previous frame's rows
        self.prev_up = torch.full((self.n_max, 32), -1, dtype=i32, device=d)
        self.warm = False
        self.logit = torch.empty(self.n_max * self.cap, dtype=f32, device=d)
        self.set_params(params)

    def set_params(self, params):
        self.params = np.asarray(params, np.float32).copy()
        self.packed = pack(self.params, self.device)

    def candidates(self, pos, rad, n):
        st = fl.stream()
        self.pos[:n].copy_(pos[:n])
        fl.call("fos_grid_build", self.ctx.handle, fl.ptr(self.pos), n, self.G, self.L,
                fl.ptr(self.cc), fl.ptr(self.cs), fl.ptr(self.ci), st)
        fl.call("fos_pred_candidates", fl.ptr(self.pos), fl.ptr(rad), n, self.G, self.L,
                self.mult, self.cap, fl.ptr(self.cs), fl.ptr(self.cc), fl.ptr(self.ci),
                fl.ptr(self.cand), fl.ptr(self.ncand), fl.ptr(self.nup), fl.ptr(self.rel),
                fl.ptr(self.trunc), st)
        fl.call("fos_pred_pairs", self.ctx.handle, fl.ptr(self.nup), n, fl.ptr(self.pair_off),
                fl.ptr(self.pair_i), st)

    def npairs(self, n):
        """Number of scored pairs of the last candidates() call (syncs)."""
        return int(self.pair_off[n].item())

    def features(self, rad, n, r_mean=None):
        if r_mean is None:
            r_mean = float(rad[:n].mean().item())
        fl.call("fos_pred_features", fl.ptr(rad), n, float(r_mean), CLIP, fl.ptr(self.cand),
                fl.ptr(self.ncand), fl.ptr(self.rel), fl.ptr(self.pair_off), fl.ptr(self.pair_i),
                self.cap, n * self.cap, fl.ptr(self.prev_up), fl.ptr(self.feat), fl.stream())

    def logits(self, n):
        pk = self.packed
        fl.call("fos_pred_mlp", fl.ptr(self.feat), n * self.cap, fl.ptr(pk["w1p"]),
                fl.ptr(pk["w2p"]), fl.ptr(pk["b1"]), fl.ptr(pk["b2"]), fl.ptr(pk["w3"]),
                pk["b3"], fl.ptr(self.logit), fl.ptr(self.pair_off[n:]), fl.stream())

    def fsc(self, n, fsc, face_ids=None, face_counts=None):
        fl.call("fos_pred_fsc", fl.ptr(self.logit), fl.ptr(self.cand), fl.ptr(self.pair_off), n,
                self.cap, fl.ptr(fsc), fl.ptr(face_ids), fl.ptr(face_counts), fl.stream())

    def predict(self, pos, rad, n, fsc, face_ids=None, face_counts=None, r_mean=None):
        """Whole pipeline for one frame: fsc[n] (and face rows when given;
        they become the next frame's previous accepted pairs)."""
        self.candidates(pos, rad, n)
        self.features(rad, n, r_mean)
        self.logits(n)
        self.fsc(n, fsc, face_ids, face_counts)
        if face_ids is not None:
            self.prev_up[:n].copy_(face_ids[:n])
            self.warm = True

    def remember_jfa(self, neighbor_list, neighbor_count, n):
        """After a JFA frame: its accepted pairs are the next frame's
        previous ones."""
        fl.call("fos_pred_prev_rows", fl.ptr(neighbor_list), fl.ptr(neighbor_count), n,
                fl.ptr(self.prev_up), fl.stream())
        self.warm = True

    def labels(self, n, neighbor_list, neighbor_count, out):
        fl.call("fos_pred_labels", fl.ptr(self.cand), fl.ptr(self.pair_off), fl.ptr(self.pair_i),
                n, self.cap, n * self.cap, fl.ptr(neighbor_list), fl.ptr(neighbor_count),
                fl.ptr(out), fl.stream())


def fsc_error(pred, ref):
    """(sum |pred - ref| / sum ref, |sum pred - sum ref| / sum ref) over a frame."""
    p, r = pred.double(), ref.double()
    tot = float(r.sum().item())
    if tot == 0:
        return 0.0, 0.0
    return float((p - r).abs().sum().item()) / tot, abs(float(p.sum().item()) - tot) / tot


TRAIN_SCRATCH = 64 * (NPARAM + 1)   # fos.h FOS_MLP_TRAIN_SCRATCH


class Trainer:
    """Adam on the fp32 parameters with the hand-written MFMA training step
    (BF16 GEMM operands, FP32 accumulation and optimizer state)."""

    def __init__(self, params, device="cuda", batch=4096, lr=2e-3, pos_weight=1.0):
        d = torch.device(device)
        self.p = torch.from_numpy(np.asarray(params, np.float32).copy()).to(d)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.B, self.lr, self.pos_weight = int(batch), float(lr), float(pos_weight)
        self.scratch = torch.empty(TRAIN_SCRATCH, dtype=torch.float32, device=d)
        self.loss = torch.zeros(1, dtype=torch.float32, device=d)
        self.t = 0

    def epoch(self, feat, labels, valid_idx, gen=None):
        """One pass over valid_idx (int64 device tensor of sample rows) in
        shuffled batches; returns the mean batch loss."""
        perm = valid_idx[torch.randperm(valid_idx.numel(), device=valid_idx.device,
                                        generator=gen)].to(torch.int32).contiguous()
        nb = perm.numel() // self.B
        self.loss.zero_()
        base, st = perm.data_ptr(), fl.stream()
        for b in range(nb):
            self.t += 1
            fl.call("fos_mlp_train_step", fl.ptr(self.p), fl.ptr(self.g), fl.ptr(self.m),
                    fl.ptr(self.v), self.t, self.lr, fl.ptr(feat), fl.ptr(labels),
                    ctypes.c_void_p(base + 4 * b * self.B), self.B, self.pos_weight,
                    fl.ptr(self.scratch), fl.ptr(self.loss), st)
        return float(self.loss.item()) / max(1, nb)

    def params(self):
        return self.p.cpu().numpy().copy()
