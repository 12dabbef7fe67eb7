"""Frame records of the packed detector formats (host only, no GPU): the
records xpcs_packed_records builds put every word in the output frame the
reference reader gives it (the oracle's restatement of Ufxc::Ufxc and
Rigaku::Rigaku, src/xpcs/io/ufxc.cpp:59-103, src/xpcs/io/rigaku.cpp:128-255)."""
import numpy as np
import pytest

import synth
import xpcs_mt


def word_frames(records, n, keep=None):
    """Output frame of every word from the records (-1: in no record); words
    where `keep` is False (masked pixels) are reported as -1."""
    fw, cnt, fr = records
    out = np.full(n, -1, np.int64)
    for b, c, f in zip(fw, cnt, fr):
        assert np.all(out[b:b + c] == -1), "a word in two records"
        out[b:b + c] = f
    if keep is not None:
        out[~keep] = -1
    return out


def rigaku_mapped_words(w, W, H, mapped):
    import oracle
    pix = oracle._raster((w >> np.uint64(16)) & np.uint64(0xFFFFF), W, H)
    return mapped[pix]


def counts(W, H, F, mu, seed):
    rng = np.random.default_rng(seed)
    return rng.poisson(mu, (F, W * H)).astype(np.int64)


def maps(W, H):
    return synth.annular_qmaps(W, H, 4, 2)


def ufxc_expected(words, F, stride, avg):
    import oracle
    ff, _, _ = oracle.ufxc_decode(words, 1, 1)
    R = stride * avg
    want = np.full(ff.size, -1, np.int64)
    for f in range(F):
        for j in range(avg):
            want[ff == f * R + j * stride] = f
    return want


@pytest.mark.parametrize("stride,avg,frame_from", [(1, 1, 0), (2, 1, 0), (1, 3, 0), (2, 2, 0), (1, 1, 7)])
def test_ufxc_records(stride, avg, frame_from):
    """Counter wrap (2040 -> 0), a frame split around the next one; frame_from
    changes nothing (Ufxc::SkipFrames does nothing)."""
    W, H, Fraw = 16, 12, 60
    c = np.minimum(counts(W, H, Fraw, 0.3, 5), 3)
    w = synth.ufxc_words(c, W, H, counter0=2040, split=(3, 20))
    dq, sq = maps(W, H)
    F = Fraw // (stride * avg)
    rec = xpcs_mt.packed_records(w, xpcs_mt.FMT_UFXC, W, H, F, dq, sq, stride, avg, frame_from)
    got = word_frames(rec, w.size)
    assert np.array_equal(got, ufxc_expected(w, F, stride, avg))
    assert (got >= 0).sum() > 0


def test_ufxc_frames_past_the_data_are_empty():
    W, H = 8, 8
    c = np.minimum(counts(W, H, 10, 0.5, 6), 3)
    w = synth.ufxc_words(c, W, H)
    dq, sq = maps(W, H)
    rec = xpcs_mt.packed_records(w, xpcs_mt.FMT_UFXC, W, H, 25, dq, sq)
    assert np.array_equal(word_frames(rec, w.size), ufxc_expected(w, 25, 1, 1))


def rigaku_case(W=16, H=12, F=40, seed=7, **kw):
    c = counts(W, H, F, 0.2, seed)
    return c, synth.rigaku_words(c, W, H, **kw)


@pytest.mark.parametrize("stride,avg,start,F", [(1, 1, 0, 40), (2, 1, 0, 15), (1, 3, 0, 12), (2, 2, 1, 8),
                                                 (1, 1, 5, 30), (1, 1, 0, 17)])
def test_rigaku_records(stride, avg, start, F):
    """Frame gaps (file frames without words make no output frame), stride and
    averaging state, skipped leading frames, the output-frame limit."""
    import oracle
    W, H = 16, 12
    c = counts(W, H, 40, 0.2, 8)
    c[[4, 5, 11, 30]] = 0                      # file frames without words
    w = synth.rigaku_words(c, W, H, first_frame=1)
    dq, sq = maps(W, H)
    mapped = (dq.reshape(-1) > 0) & (sq.reshape(-1) > 0)
    rec = xpcs_mt.packed_records(w, xpcs_mt.FMT_RIGAKU, W, H, F, dq, sq, stride, avg, start)
    want = oracle.rigaku_assign(w, W, H, mapped, start, F, stride, avg)
    assert np.array_equal(word_frames(rec, w.size, rigaku_mapped_words(w, W, H, mapped)), want)
    assert len(rec[0]) <= F * stride * avg + 1     # runs break only at frame changes


def test_rigaku_last_frame_needs_two_pixels():
    import oracle
    W, H = 8, 8
    dq, sq = maps(W, H)
    mapped = (dq.reshape(-1) > 0) & (sq.reshape(-1) > 0)
    px = np.flatnonzero(mapped)
    c = np.zeros((3, W * H), np.int64)
    c[0, px[:5]] = 1
    c[1, px[:3]] = 2
    for extra, kept in (([(3, px[0], 5)], False), ([(3, px[0], 5), (3, px[1], 1)], True),
                        ([(3, px[0], 5), (3, px[0], 1)], False)):
        w = synth.rigaku_words(c[:2], W, H, extra=extra)
        rec = xpcs_mt.packed_records(w, xpcs_mt.FMT_RIGAKU, W, H, 10, dq, sq)
        got = word_frames(rec, w.size, rigaku_mapped_words(w, W, H, mapped))
        assert np.array_equal(got, oracle.rigaku_assign(w, W, H, mapped, 0, 10))
        assert (got[-len(extra):] >= 0).all() == kept


def test_rigaku_pixel_out_of_range_is_left_to_the_device():
    """Pixels are decoded where the words are used (the transpose flags them);
    the host pass only assigns frames."""
    W, H = 8, 8
    dq, sq = maps(W, H)
    w = np.array([(1 << 40) | (2000 << 16) | 1, (2 << 40) | (3 << 16) | 1], np.uint64)   # raster 250 of 64
    rec = xpcs_mt.packed_records(w, xpcs_mt.FMT_RIGAKU, W, H, 5, dq, sq)
    assert word_frames(rec, w.size)[0] == 0


def test_packed_sizes_and_format():
    dq, sq = maps(8, 8)
    with pytest.raises(xpcs_mt.XpcsError):
        xpcs_mt.packed_records(np.zeros(3, np.uint8), xpcs_mt.FMT_UFXC, 8, 8, 5, dq, sq)
    with pytest.raises(xpcs_mt.XpcsError):
        xpcs_mt.packed_records(np.zeros(12, np.uint8), xpcs_mt.FMT_RIGAKU, 8, 8, 5, dq, sq)
    with pytest.raises(xpcs_mt.XpcsError):
        xpcs_mt.packed_records(np.zeros(8, np.uint8), 9, 8, 8, 5, dq, sq)
