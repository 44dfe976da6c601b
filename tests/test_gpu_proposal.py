"""The proposal network's HIP kernels against the CPU oracle, and the gated
search against the plain search.

Bars: pooling bit-exact; first-layer activations within one BF16 ulp (the GPU
sums in FP32 in matrix-core order, the oracle in float64, so a value near a
rounding boundary may round the other way), deeper ones identical in >= 99 %
of elements and within 2^-7 relative + 1 % of the layer's RMS;
logits within 1e-2; the gradient within 3 % relative L2 error of the oracle's
(same argmax positions); the gated search with every block flagged equals the
plain device-statistics search hit for hit; the gated batch equals the per-call
gated search on every field (and the plain batch with every block flagged).
"""

import numpy as np
import pytest

import seticore_amd as sc
from seticore_amd import proposal

pytestmark = pytest.mark.gpu

T, FOFF, TSAMP = 16, -2.79e-6, 18.25


def _coarse_channel(torch, N, seed):
    rng = np.random.default_rng(seed)
    recipe = proposal.coarse_channel_recipe(rng, N, num_tones=10, num_rfi=2)
    x = torch.empty((T, N), dtype=torch.float32, device="cuda")
    proposal.synth(x, seed, recipe)
    return x


def test_pool_bit_exact(gpu, oracle):
    import torch
    from oracle import proposal_oracle as P
    x = _coarse_channel(torch, 1 << 16, 1)
    stats = torch.tensor([16.3, 4.1], dtype=torch.float32, device="cuda")
    got = sc.ProposalNet.pool(x, stats).float().cpu().numpy()
    want = P.pool(x.cpu().numpy(), np.float32(16.3), np.float32(4.1))
    np.testing.assert_array_equal(got, want)


def _tiles_and_params(torch, B, seed):
    x = _coarse_channel(torch, B * 4096, seed)
    col = x.sum(0)
    stats = torch.tensor([float(col.median()), float(col.std())], dtype=torch.float32,
                         device="cuda")
    tiles = sc.ProposalNet.pool(x, stats)
    net = sc.ProposalNet(max_batch=B, seed=seed)
    p = net.get_params()
    # Small positive biases so that every layer has active units.
    rng = np.random.default_rng(seed)
    from oracle import proposal_oracle as P
    p[P.OFF_B1:P.OFF_W2] = rng.uniform(0, 0.1, P.C1)
    p[P.OFF_B2:P.OFF_W3] = rng.uniform(0, 0.1, P.C2)
    p[P.OFF_B3:P.OFF_WH] = rng.uniform(0, 0.1, P.C3)
    net.set_params(p)
    return net, tiles, p


def _bf16_ulps(a, b):
    ua = a.astype(np.float32).view(np.int32) >> 16
    ub = b.astype(np.float32).view(np.int32) >> 16
    return np.abs(ua.astype(np.int64) - ub.astype(np.int64))


def test_forward_matches_oracle(gpu):
    import torch
    from oracle import proposal_oracle as P
    B = 6
    net, tiles, p = _tiles_and_params(torch, B, 11)
    logits = net.forward(tiles).cpu().numpy()
    h1, h2, h3, _, _ = net.activations(B)
    zl, _, (r1, r2, r3) = P.forward(tiles.float().cpu().numpy(), p, bf16=True)
    # h1 comes from identical inputs: at most one ulp apart. Deeper layers
    # inherit those one-ulp differences of their inputs, which matter most
    # where a sum cancels to a small value: bound them by the layer's scale.
    g1 = h1.float().cpu().numpy()
    assert _bf16_ulps(g1, r1).max() <= 1
    for got, want in ((h1, r1), (h2, r2), (h3, r3)):
        g = got.float().cpu().numpy()
        assert (_bf16_ulps(g, want) == 0).mean() >= 0.99
        scale = np.sqrt(np.mean(want ** 2))
        assert np.all(np.abs(g - want) <= 2.0 ** -7 * np.abs(want) + 1e-2 * scale)
        assert (want > 0).mean() > 0.05  # the check covers active units
    np.testing.assert_allclose(logits, zl, rtol=1e-2, atol=1e-2)


def test_gradient_matches_oracle(gpu):
    import torch
    from oracle import proposal_oracle as P
    B = 8
    net, tiles, p = _tiles_and_params(torch, B, 12)
    labels = torch.tensor([1, 0, 0, 1, 0, 0, 0, 1], dtype=torch.float32, device="cuda")
    grad, loss = net.gradient(tiles, labels, pos_weight=4.0)
    g = grad.cpu().numpy().astype(np.float64)
    _, _, _, zl, arg = net.activations(B)
    ref_loss, ref = P.gradient(tiles.float().cpu().numpy(), labels.cpu().numpy(), p, 4.0)
    _, ref_arg, _ = P.forward(tiles.float().cpu().numpy(), p, bf16=True)
    np.testing.assert_array_equal(arg.cpu().numpy(), ref_arg)
    assert abs(float(loss.item()) - ref_loss) <= 1e-3 * max(1.0, ref_loss)
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print("gradient relative L2 error %.2e" % err)
    assert err < 3e-2
    # Every layer carries gradient.
    for lo, hi in [(P.OFF_W1, P.OFF_B1), (P.OFF_W2, P.OFF_B2), (P.OFF_W3, P.OFF_B3),
                   (P.OFF_WH, P.OFF_BH)]:
        assert np.abs(ref[lo:hi]).max() > 0


def test_gated_search_with_every_block_flagged_equals_plain(gpu):
    import torch
    N = 1 << 18
    x = _coarse_channel(torch, N, 21)
    meta = sc.FilterbankMetadata(coarse_channel_size=N, foff=FOFF)
    buf = sc.FilterbankBuffer(T, N, data=x)
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    plain = []
    dd.search(buf, meta, -1, 0, 10.0, 0.0, 10.0, plain)
    st_plain = dd.state()
    net = sc.ProposalNet(max_batch=64, seed=5)
    gated = []
    flagged = dd.search_proposed(buf, net, meta, -1, 0, 10.0, 0.0, 10.0, -1e30, gated)
    assert flagged == N // 4096
    st = dd.state()
    np.testing.assert_array_equal(st["top_sums"], st_plain["top_sums"])
    assert [h.key() + (h.snr,) for h in gated] == [h.key() + (h.snr,) for h in plain]
    assert len(plain) > 0
    # Nothing flagged: nothing searched, no hits.
    none = []
    assert dd.search_proposed(buf, net, meta, -1, 0, 10.0, 0.0, 10.0, 1e30, none) == 0
    assert none == []


def _fields(hits):
    return [(h.index, h.drift_steps, h.drift_rate, h.snr, h.power, h.frequency, h.beam,
             h.coarse_channel, h.num_timesteps) for h in hits]


def test_gated_batch_equals_per_call_gated_search(gpu):
    # The gated batch (three streams: trees; each coarse channel's
    # statistics, network and tile list beside the previous channel's tree;
    # merges) against the per-call gated search of each coarse channel: the
    # same tiles searched, the same blocks flagged, every hit field equal -
    # FP32 and BF16 storage mixed in one batch, two runs in a row. Below
    # every logit it is the plain batch, hit for hit.
    import torch
    from seticore_amd import batch as B
    N, n_cc = 1 << 18, 6
    nblk = N // 4096
    net = sc.ProposalNet(max_batch=nblk, seed=5)
    xs = [_coarse_channel(torch, N, 40 + i) for i in range(n_cc)]
    dtypes = [sc.SC_F32 if i % 2 == 0 else sc.SC_BF16 for i in range(n_cc)]
    data = [x if d == sc.SC_F32 else sc.normalize(x, sc.SC_BF16)[0] for x, d in zip(xs, dtypes)]
    meta = sc.FilterbankMetadata(coarse_channel_size=N, foff=FOFF)
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    # A threshold that flags about half of the blocks: the median logit.
    logits = []
    for x in xs:
        dd.search(sc.FilterbankBuffer(T, N, data=x), meta, -1, 0, 10.0, 0.0, 10.0, [])
        st = dd.state()
        stats = torch.tensor([st["median"], st["std"]], dtype=torch.float32, device="cuda")
        logits.append(net.forward(net.pool(x, stats)).cpu())
    thr = float(torch.cat(logits).median())
    bufs = [sc.FilterbankBuffer(T, N, data=d, dtype=t) for d, t in zip(data, dtypes)]
    inst = [B.instance(d, T, FOFF, TSAMP, 10.0, 0.0, 10.0, coarse_channel=i, dtype=t,
                       metadata=meta) for i, (d, t) in enumerate(zip(data, dtypes))]
    per_call = []
    for i in range(n_cc):
        out = []
        flagged = dd.search_proposed(bufs[i], net, meta, -1, i, 10.0, 0.0, 10.0, thr, out)
        per_call.append((flagged, out))
    flagged_total = sum(f for f, _ in per_call)
    print("flagged %d of %d blocks, %d hits" % (flagged_total, n_cc * nblk,
                                                sum(len(o) for _, o in per_call)))
    assert 0 < flagged_total < n_cc * nblk
    gb = B.Batch(inst, proposal=net, logit_threshold=thr)
    for _ in range(2):
        gb.run()
        for i, (flagged, out) in enumerate(per_call):
            assert gb.flagged(i)[1] == flagged, i
            assert _fields(gb.hits(i)) == _fields(out), i
    assert any(out for _, out in per_call)
    # Every block flagged: the plain batch's hits.
    everything = B.Batch(inst, proposal=net, logit_threshold=-1e30)
    plain = B.Batch(inst)
    everything.run()
    plain.run()
    for i in range(n_cc):
        assert everything.flagged(i)[1] == nblk
        assert _fields(everything.hits(i)) == _fields(plain.hits(i)), i
    with pytest.raises(sc.SeticoreError):
        plain.flagged(0)


def test_gated_batch_of_mixed_widths(gpu):
    # Two widths interleaved (four search handles, each instance waiting for
    # the tail of the previous instance on its own handle): every instance
    # equals its per-call gated search, hits and flagged blocks.
    import torch
    from seticore_amd import batch as B
    widths = [1 << 17, 1 << 16, 1 << 17, 1 << 16, 1 << 16, 1 << 17, 1 << 17]
    net = sc.ProposalNet(max_batch=max(widths) // 4096, seed=9)
    xs = [_coarse_channel(torch, N, 60 + i) for i, N in enumerate(widths)]
    logits = []
    for x, N in zip(xs, widths):
        dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
        dd.set_stats_mode(sc.SC_STATS_DEVICE)
        meta = sc.FilterbankMetadata(coarse_channel_size=N, foff=FOFF)
        dd.search(sc.FilterbankBuffer(T, N, data=x), meta, -1, 0, 10.0, 0.0, 10.0, [])
        st = dd.state()
        stats = torch.tensor([st["median"], st["std"]], dtype=torch.float32, device="cuda")
        logits.append(net.forward(net.pool(x, stats)).cpu())
    thr = float(torch.cat(logits).median())  # about half of the blocks flagged
    per_call, inst = [], []
    for i, (x, N) in enumerate(zip(xs, widths)):
        meta = sc.FilterbankMetadata(coarse_channel_size=N, foff=FOFF)
        dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
        dd.set_stats_mode(sc.SC_STATS_DEVICE)
        out = []
        flagged = dd.search_proposed(sc.FilterbankBuffer(T, N, data=x), net, meta, -1, i, 10.0,
                                     0.0, 10.0, thr, out)
        per_call.append((flagged, out))
        inst.append(B.instance(x, T, FOFF, TSAMP, 10.0, 0.0, 10.0, coarse_channel=i,
                               metadata=meta))
    gb = B.Batch(inst, proposal=net, logit_threshold=thr)
    assert gb.info()["shapes"] == 2
    gb.run()
    for i, (flagged, out) in enumerate(per_call):
        assert gb.flagged(i)[1] == flagged, i
        assert _fields(gb.hits(i)) == _fields(out), i
    total = sum(f for f, _ in per_call)
    print("flagged %d of %d blocks" % (total, sum(widths) // 4096))
    assert 0 < total < sum(widths) // 4096
    assert any(out for _, out in per_call)


def test_trained_network_beats_the_brightest_pixel_heuristic(gpu):
    # What training buys, against the non-learned baseline: on 80 held-out
    # coarse channels, at the threshold that keeps 90 % of the plain search's
    # hit blocks, the trained network flags fewer blocks than the brightest
    # pooled pixel does (and than the same network untrained). The bar is the
    # claim itself - strictly fewer blocks than the heuristic - with the
    # margin printed; measured: trained 0.748, heuristic 0.783, untrained
    # 0.909 (deterministic: seeded data, fixed-order gradient sums).
    from seticore_amd import proposal as P
    N = 1 << 18
    train = P.make_dataset(640, N, seed=31)
    val = P.make_dataset(80, N, seed=32)
    net = sc.ProposalNet(max_batch=256, seed=33)
    z0 = P.scores(net, val.tiles)
    hist = P.train(net, train)
    z1 = P.scores(net, val.tiles)
    zh = P.heuristic_scores(val.tiles)
    flags = []
    for z in (z0, z1, zh):
        thr = P.threshold_for_recall(z, val.labels, 0.9)
        flags.append(float((z > thr).float().mean()))
    print("loss %.3f -> %.3f; blocks flagged at 90%% block recall: untrained %.3f, trained "
          "%.3f, brightest pixel %.3f (margin %.3f)" % (hist[0], hist[-1], flags[0], flags[1],
                                                          flags[2], flags[2] - flags[1]))
    assert hist[-1] < hist[0]
    assert flags[1] < flags[2]
    assert flags[1] < flags[0]


@pytest.mark.parametrize("dtype", [sc.SC_BF16, sc.SC_F16])
def test_gated_batch_runs_share_a_tree_launch(gpu, dtype):
    # Eleven BF16 or FP16 coarse channels of one width: runs of eight and
    # three share each front launch and one launch of the fused kernel over
    # their tile lists (each channel its own list, about half of its blocks
    # flagged). Every instance equals its per-call gated search, hits and
    # flagged blocks, over two runs in a row.
    import torch
    from seticore_amd import batch as B
    N, n_cc = 1 << 17, 11
    nblk = N // 4096
    net = sc.ProposalNet(max_batch=nblk, seed=11)
    xs = [_coarse_channel(torch, N, 80 + i) for i in range(n_cc)]
    data = [sc.normalize(x, dtype)[0] for x in xs]
    meta = sc.FilterbankMetadata(coarse_channel_size=N, foff=FOFF)
    dd = sc.Dedopplerer(T, N, FOFF, TSAMP)
    dd.set_stats_mode(sc.SC_STATS_DEVICE)
    logits = []
    for x in data:
        dd.search(sc.FilterbankBuffer(T, N, data=x, dtype=dtype), meta, -1, 0, 10.0, 0.0,
                  10.0, [])
        st = dd.state()
        stats = torch.tensor([st["median"], st["std"]], dtype=torch.float32, device="cuda")
        logits.append(net.forward(net.pool(x, stats)).cpu())
    with pytest.raises(ValueError):  # the tensor's sample type is not FP32
        net.pool(data[0], stats, sc.SC_F32)
    thr = float(torch.cat(logits).median())
    per_call, inst = [], []
    for i, x in enumerate(data):
        out = []
        flagged = dd.search_proposed(sc.FilterbankBuffer(T, N, data=x, dtype=dtype), net,
                                     meta, -1, i, 10.0, 0.0, 10.0, thr, out)
        per_call.append((flagged, out))
        inst.append(B.instance(x, T, FOFF, TSAMP, 10.0, 0.0, 10.0, coarse_channel=i,
                               dtype=dtype, metadata=meta))
    total = sum(f for f, _ in per_call)
    print("flagged %d of %d blocks, %d hits" % (total, n_cc * nblk,
                                                sum(len(o) for _, o in per_call)))
    assert 0 < total < n_cc * nblk
    assert any(out for _, out in per_call)
    gb = B.Batch(inst, proposal=net, logit_threshold=thr)
    for _ in range(2):
        gb.run()
        for i, (flagged, out) in enumerate(per_call):
            assert gb.flagged(i)[1] == flagged, i
            assert _fields(gb.hits(i)) == _fields(out), i
    # Nothing flagged in any channel: the set launch's items all exit.
    none = B.Batch(inst, proposal=net, logit_threshold=1e30)
    none.run()
    for i in range(n_cc):
        assert none.flagged(i) == (0, 0)
        assert none.hits(i) == []
