"""GL codebook training and VQ coding with the nearest-codeword search on the
GPU (glp_vq_train_device, glp_vq_codes_device, glp_vq_assign_device) against
the oracle's GLVQ restatement (glVQ.cpp:26-165, glPack.cpp:195-219): the same
codebook bit for bit, the same iteration count, the same codes -- including
codebooks with repeated codewords and squared distances one or two units in
the last place apart, where the first of two codewords at equal rounded
distance must win, and the 5,000-sample x 4,096-site GL set of the MH update
benchmark."""
import sys

import numpy as np
import pytest

from glphase_amd import host, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gls(N, M, seed, depth=4.0):
    return synth.make_dataset(N, M, depth=depth, seed=seed)["gls"]


@pytest.mark.parametrize("N,M,codes,training", [(40, 300, 256, 0), (200, 500, 16, 0),
                                                (300, 1000, 256, 1 << 15), (64, 256, 4, 0)])
def test_vq_train_device_matches_oracle(orc, N, M, codes, training):
    say(f"vq train N={N} M={M} codes={codes} training={training}")
    gls = _gls(N, M, seed=N + M)
    cb, it = host.vq_train(gls, codes, key=5, training_size=training, device=True)
    rc, cbo, ito = orc.vq_train(gls, codes, 5, training)
    assert rc == 0
    assert it == ito
    assert cb.tobytes() == cbo.tobytes()
    got = host.vq_codes(gls, cb, device=True)
    assert (got == orc.vq_codes(gls, cb)).all()


def test_vq_codes_device_ties_take_first(orc):
    say("vq codes with repeated codewords")
    rng = np.random.default_rng(4)
    base = rng.random((40, 2)).astype(np.float32) * np.float32(0.5)
    cb = np.concatenate([base, base[::-1], base[:8]]).astype(np.float32)  # every codeword twice+
    # GL trios whose normalised (RR, Het) land exactly on codewords, and
    # points half-way between two codewords
    aa = np.float32(1) - cb[:, 0] - cb[:, 1]
    on = np.stack([cb[:, 0], cb[:, 1], aa], 1)
    mid = (cb[:40] + cb[40:80]) / np.float32(2)
    mid = np.stack([mid[:, 0], mid[:, 1], np.float32(1) - mid[:, 0] - mid[:, 1]], 1)
    rnd = rng.random((5000, 3)).astype(np.float32)
    gls = np.concatenate([on, mid, rnd]).astype(np.float32)
    got = host.vq_codes(gls, cb, device=True)
    want = orc.vq_codes(gls, cb)
    assert (got == want).all()
    assert (got == host.vq_codes(gls, cb)).all()


def test_vq_assign_device_on_pairs(orc):
    say("vq assign on normalised pairs")
    import torch
    from glphase_amd._lib import c_f32p, check, lib, ptr
    gls = _gls(100, 400, seed=2).reshape(-1, 3)
    cb, _ = host.vq_train(gls, 64, key=1)
    s = gls.sum(1, dtype=np.float32)
    pts = np.stack([gls[:, 0] / s, gls[:, 1] / s], 1).astype(np.float32)
    dev = torch.from_numpy(pts).cuda()
    out = torch.zeros(pts.shape[0], dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().glp_vq_assign_device(dev.data_ptr(), pts.shape[0], 0, ptr(cb, c_f32p), 64,
                                     out.data_ptr(), stream), "glp_vq_assign_device")
    torch.cuda.synchronize()
    want = np.array([orc.vq_find(cb, float(r), float(h)) for r, h in pts[:4000]], np.uint8)
    assert (out.cpu().numpy()[:4000] == want).all()


def _near_ties(count, seed=1):
    """Codeword pairs (a, 0), (b1, b2) whose squared distances from (0, 0)
    differ in the last bits: s1 * (1 - 2^-50) <= s2 < s1.  Returns the pairs
    whose double roots are equal (the first codeword must win) and those
    whose roots differ (the nearer one must win)."""
    rng = np.random.default_rng(seed)
    eq, lt = [], []
    while len(eq) < count or len(lt) < count:
        a = np.float32(rng.uniform(0.3, 0.9))
        b1 = np.nextafter(a, np.float32(0))
        s1 = np.float64(a) ** 2
        b2c = np.float32(np.sqrt(s1 - np.float64(b1) ** 2))
        for b2 in (b2c, np.nextafter(b2c, np.float32(0)), np.nextafter(b2c, np.float32(1))):
            s2 = np.float64(b1) ** 2 + np.float64(b2) ** 2
            if s1 * (1 - 2.0**-50) <= s2 < s1:
                (eq if np.sqrt(s2) == np.sqrt(s1) else lt).append((a, b1, b2))
    return eq[:count], lt[:count]


def test_vq_assign_device_near_ties(orc):
    say("vq assign on near-tie distances")
    import torch
    from glphase_amd._lib import c_f32p, check, lib, ptr
    eq, lt = _near_ties(24)
    pt = torch.zeros((1, 2), dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for kind, cases in (("equal roots", eq), ("distinct roots", lt)):
        for a, b1, b2 in cases:
            for cb in (np.array([[a, 0], [b1, b2]], np.float32),
                       np.array([[b1, b2], [a, 0]], np.float32)):
                check(lib().glp_vq_assign_device(pt.data_ptr(), 1, 0, ptr(cb, c_f32p), 2,
                                                 out.data_ptr(), stream), "glp_vq_assign_device")
                want = orc.vq_find(cb, 0.0, 0.0)
                assert int(out.cpu()[0]) == want, (kind, cb.tolist())
            if kind == "equal roots":
                assert want == 0  # the later codeword lost the tie


def test_vq_assign_device_near_ties_in_groups(orc):
    # the search takes codewords 1.. in branch-free groups of 4 (then a
    # scalar tail) and redoes a group when a lane of the wavefront met a near
    # tie in it: near-tie pairs placed inside a group (2/3, 9/10), across a
    # group boundary (4/5, 8/9) and in the tail (13/14) of a 16-word book
    # padded with far codewords, in both orders, with 1,024 points per launch
    # so that lanes at the tie (the origin) share wavefronts with lanes that
    # have none
    say("vq assign: near ties inside and across 4-codeword groups")
    import torch
    from glphase_amd._lib import c_f32p, check, lib, ptr
    eq, lt = _near_ties(6, seed=5)
    rng = np.random.default_rng(9)
    n = 1024
    pts = rng.random((n, 2)).astype(np.float32)
    pts[::37] = 0.0
    dev = torch.from_numpy(pts).cuda()
    out = torch.zeros(n, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    far = np.array([[3.0 + k, 3.0 + 0.5 * k] for k in range(16)], np.float32)
    ties = 0
    for a, b1, b2 in eq + lt:
        for i, j in ((2, 3), (9, 10), (4, 5), (8, 9), (13, 14)):
            for first, second in (((a, 0), (b1, b2)), ((b1, b2), (a, 0))):
                cb = far.copy()
                cb[i], cb[j] = first, second
                check(lib().glp_vq_assign_device(dev.data_ptr(), n, 0, ptr(cb, c_f32p), 16,
                                                 out.data_ptr(), stream), "glp_vq_assign_device")
                got = out.cpu().numpy()
                want = np.array([orc.vq_find(cb, float(r), float(h)) for r, h in pts], np.uint8)
                assert (got == want).all(), (i, j, cb[i].tolist(), cb[j].tolist())
                assert set(got[::37]) <= {i, j}
                ties += int(got[0] == i)
    assert ties > 0


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
@pytest.mark.parametrize("N,M,s0,s1", [(1, 1, 0, 1), (70, 300, 0, 70), (130, 1000, 7, 100),
                                       (5000, 4096, 0, 5000)])
def test_pack_codes_device_matches_oracle(orc, bits, N, M, s0, s1):
    # GLPack::GetPackedGLs on the device, word for word: MSB-first codes,
    # site counts that leave a partial last word (padded with the code of
    # GL 0, glPack.cpp:66-70), sample ranges, 1/2/4/8-bit codes
    say(f"pack codes on the device bits={bits} N={N} M={M} [{s0},{s1})")
    rng = np.random.default_rng(N + M + bits)
    codes = rng.integers(0, 1 << bits, (N, M)).astype(np.uint8)
    got = host.pack_codes(codes, bits, s0, s1, device=True)
    want = orc.pack_codes(codes, bits, s0, s1)
    assert got.shape == want.shape and (got == want).all()
    assert (got == host.pack_codes(codes, bits, s0, s1)).all()


def test_vq_full_benchmark_gls(orc):
    # the GL set of the MH update benchmark: 5,000 samples x 4,096 sites
    say("vq 5000 x 4096 GLs")
    gls = _gls(5000, 4096, seed=0)
    cb, it = host.vq_train(gls, 256, key=0, training_size=1 << 16, device=True)
    cbh, ith = host.vq_train(gls, 256, key=0, training_size=1 << 16)
    assert it == ith and cb.tobytes() == cbh.tobytes()
    codes = host.vq_codes(gls, cb, device=True)
    assert (codes == host.vq_codes(gls, cb)).all()
    pick = np.random.default_rng(0).integers(0, 5000 * 4096, 200000)
    flat = gls.reshape(-1, 3)
    assert (codes.reshape(-1)[pick] == orc.vq_codes(flat[pick], cb)).all()


def test_vq_device_rejects_bad_arguments():
    from glphase_amd._lib import GlpError
    gls = np.ones((10, 3), np.float32)
    with pytest.raises(GlpError):
        host.vq_codes(gls, np.zeros((300, 2), np.float32), device=True)
    with pytest.raises(GlpError):
        host.vq_train(gls, 2, device=True)
