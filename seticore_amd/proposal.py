"""Training, calibration and evaluation of the block proposal network.

The network (seticore_amd/csrc/pn_kernels.hip) scores 4096-channel blocks;
its labels come from the plain search itself: a block is positive when the
plain search reports a hit whose start channel lies in it. Everything runs on
the GPU through the C ABI: synthetic coarse channels (seeded), the plain
search that labels them, pooling, forward/backward passes and Adam.

Synthetic coarse channels follow the file-scale workload (16 time steps,
foff -2.79e-6 MHz, tsamp 18.25 s, exponential noise with a gain U[0.5, 2], a
cosine taper over the outer 5 %, a DC spike of 50x the mean at the centre bin)
with tones at SNR 8-50 and drift +-10 Hz/s, plus stationary narrowband RFI at
up to 1e6 x the noise and broadband bursts on ~3 % of the time steps at 100 x
the noise.
"""

import math

import numpy as np

from . import (BLOCK_CHANNELS, SC_F32, SC_STATS_DEVICE, Dedopplerer, FilterbankBuffer,
               FilterbankMetadata, ProposalNet, gen_waterfall)

T = 16
FOFF = -2.79e-6
TSAMP = 18.25
MAX_DRIFT = 10.0
SNR_THRESHOLD = 10.0


def _torch():
    import torch
    return torch


def coarse_channel_recipe(rng, N, num_tones=12, num_rfi=3, snr=(8.0, 50.0), drift=10.0,
                          burst_rate=0.03, rfi_max=1e6):
    """Random description of one coarse channel: gain, tones (start, slope,
    amplitude per unit gain), RFI lines and burst rows."""
    gain = float(rng.uniform(0.5, 2.0))
    tones = []
    for _ in range(num_tones):
        d = float(rng.uniform(-drift, drift))
        slope = d * TSAMP / (FOFF * 1e6)
        amp = float(rng.uniform(*snr)) / math.sqrt(T)
        tones.append((float(rng.uniform(0.06 * N, 0.94 * N)), slope, amp))
    for _ in range(num_rfi):
        tones.append((float(rng.integers(int(0.06 * N), int(0.94 * N))), 0.0,
                      float(10 ** rng.uniform(1.0, math.log10(rfi_max)))))
    bursts = [t for t in range(T) if rng.uniform() < burst_rate]
    return gain, tones, bursts


def synth(x, seed, recipe):
    """Fills the CUDA float32 tensor x [16][N] from a recipe."""
    gain, tones, bursts = recipe
    gen_waterfall(x, T, seed=seed, gain=gain, taper_frac=0.05, dc_spike=50.0,
                  tones=[(s, sl, a * gain) for s, sl, a in tones])
    for t in bursts:
        x[t] += 100.0 * gain
    return x


def block_labels(hits, num_blocks, label_snr=0.0):
    """1 for blocks holding the start channel of a plain-search hit with
    SNR >= label_snr, else 0."""
    y = np.zeros(num_blocks, np.float32)
    for h in hits:
        if h.snr >= label_snr:
            y[h.index // BLOCK_CHANNELS] = 1.0
    return y


class Dataset:
    """Pooled tiles (BF16, on the GPU), block labels and the plain hits."""

    def __init__(self, tiles, labels, hits, recipes):
        self.tiles = tiles
        self.labels = labels
        self.hits = hits
        self.recipes = recipes


def make_dataset(num_cc, N, seed, label_snr=0.0, snr_threshold=SNR_THRESHOLD, **recipe_args):
    """num_cc synthetic coarse channels of N channels: each is searched with
    the plain search (device statistics, SNR threshold `snr_threshold`) for its
    labels (hits with SNR >= label_snr), then pooled."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    nblk = N // BLOCK_CHANNELS
    tiles = torch.empty((num_cc * nblk, T, 1024), dtype=torch.bfloat16, device="cuda")
    labels = np.zeros(num_cc * nblk, np.float32)
    x = torch.empty((T, N), dtype=torch.float32, device="cuda")
    dd = Dedopplerer(T, N, FOFF, TSAMP)
    dd.set_stats_mode(SC_STATS_DEVICE)
    meta = FilterbankMetadata(coarse_channel_size=N, foff=FOFF)
    all_hits, recipes = [], []
    for cc in range(num_cc):
        recipe = coarse_channel_recipe(rng, N, **recipe_args)
        synth(x, seed * 100003 + cc, recipe)
        hits = []
        dd.search(FilterbankBuffer(T, N, data=x), meta, -1, cc, MAX_DRIFT, 0.0, snr_threshold,
                  hits)
        st = dd.state()
        stats = torch.tensor([st["median"], st["std"]], dtype=torch.float32, device="cuda")
        ProposalNet.pool(x, stats, SC_F32, out=tiles[cc * nblk:(cc + 1) * nblk])
        labels[cc * nblk:(cc + 1) * nblk] = block_labels(hits, nblk, label_snr)
        all_hits.append(hits)
        recipes.append(recipe)
    torch.cuda.synchronize()
    return Dataset(tiles, torch.from_numpy(labels).cuda(), all_hits, recipes)


def train(net, data, epochs=20, batch=256, lr=1e-2, pos_weight=1.0, seed=0, log=None):
    """Adam on shuffled mini-batches; returns the mean loss of each epoch.
    Defaults from a sweep on held-out coarse channels (lr 3e-4 .. 1e-2, 20 or
    80 epochs, unweighted or class-balanced loss): lr 1e-2, 20 epochs and an
    unweighted loss flag the fewest blocks at 90 % and 99 % block recall.
    pos_weight=None balances the classes instead."""
    torch = _torch()
    n = data.tiles.shape[0]
    batch = min(batch, net.max_batch, n)
    if pos_weight is None:
        pos = float(data.labels.sum().item())
        pos_weight = min(50.0, max(1.0, (n - pos) / max(pos, 1.0)))
    g = torch.Generator().manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    history = []
    steps_per_epoch = max(1, n // batch)
    total_steps = epochs * steps_per_epoch
    step = 0
    for ep in range(epochs):
        order = torch.randperm(n, generator=g).cuda()
        total = torch.zeros(1, dtype=torch.float32, device="cuda")
        for i in range(0, steps_per_epoch * batch, batch):
            idx = order[i:i + batch]
            tiles = data.tiles.index_select(0, idx).contiguous()
            labels = data.labels.index_select(0, idx).contiguous()
            # Cosine decay of the learning rate to 5 % of its start.
            rate = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * step / total_steps)))
            net.train_step(tiles, labels, rate, pos_weight, loss=loss)
            total += loss
            step += 1
        history.append(float(total.item()) / steps_per_epoch)
        if log:
            log("epoch %d loss %.5f" % (ep, history[-1]))
    return history


def scores(net, tiles, batch=None):
    """Block logits of every tile (CUDA float32)."""
    torch = _torch()
    batch = batch or net.max_batch
    out = [net.forward(tiles[i:i + batch].contiguous()) for i in range(0, tiles.shape[0], batch)]
    return torch.cat(out)


def threshold_for_recall(logits, labels, recall):
    """Largest logit threshold that keeps at least `recall` of the positive
    blocks above it (strictly)."""
    z = logits.detach().float().cpu().numpy()
    y = labels.detach().float().cpu().numpy() > 0.5
    pos = np.sort(z[y])
    if pos.size == 0:
        return float("inf")
    k = int(math.floor((1.0 - recall) * pos.size))  # positives allowed below
    k = min(k, pos.size - 1)
    return float(np.nextafter(pos[k], -np.inf))


def heuristic_scores(tiles):
    """Non-learned baseline: the brightest pooled pixel of each block."""
    return tiles.float().amax(dim=(1, 2))
