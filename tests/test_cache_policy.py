"""Host-only: which passes use non-temporal (streaming) loads and stores of the
caller's arrays.  A plan created without a variant streams the input and the
final output, except an N-D plan whose arrays hold at most 128 MiB per side: its
outer axes re-read dst, so its input, output and the passes between them stay
cached together in the 256 MB MALL.  A pass that writes dst for a later pass to
read always stores through the caches.  stream_io 1 / 0 force streaming on / off."""
import pytest

import plan_model as M

IL = 1  # afft_interleaved
SRC, DST = 0, 1  # buffer roles of the caller's input and output


def _nt(kind, dims, batch=1, variant=None):
    _, passes = M.inspect(kind, dims, IL, batch, variant=variant)
    return [(p.in_.role, p.out.role, p.nt_io) for p in passes]


def _check(passes, streaming):
    """Streaming only where allowed: loads of src, the last store of dst.  (Row
    passes of up to 128 points and the real row passes keep cached accesses
    even when streaming is on: their accesses cover parts of cache lines.)"""
    for i, (rin, rout, nt) in enumerate(passes):
        reread = any(q[0] == rout for q in passes[i + 1:])
        if nt & 1:
            assert streaming and rin == SRC, (i, passes)
        if nt & 2:
            assert streaming and rout == DST and not reread, (i, passes)
    if streaming:
        assert any(nt for _, _, nt in passes), passes


@pytest.mark.parametrize("kind,dims,batch", [("c", (256, 256, 256), 1), ("c", (1024, 1024), 16),
                                             ("c", (64, 64, 64), 16), ("r", (256, 256, 256), 1),
                                             ("ci", (128, 256, 512), 1)])
def test_nd_arrays_within_128_mib_stay_cached(kind, dims, batch):
    p = _nt(kind, dims, batch)
    assert len(p) >= 2
    assert all(nt == 0 for _, _, nt in p), p
    if dims[-1] > 128:
        _check(_nt(kind, dims, batch, variant={"stream_io": 1}), True)


@pytest.mark.parametrize("kind,dims,batch", [("c", (4096, 4096), 2), ("c", (512, 512, 512), 1),
                                             ("r", (8192, 8192), 1)])
def test_nd_arrays_above_128_mib_stream(kind, dims, batch):
    p = _nt(kind, dims, batch)
    assert p[0][2] & 1, p  # the first pass streams the caller's input
    _check(p, True)
    _check(_nt(kind, dims, batch, variant={"stream_io": 0}), False)


def test_one_dim_four_step_streams_whatever_its_size():
    # 64 x 2^20 (512 MiB) and 4 x 2^20 (32 MiB): pass A streams src in, pass B streams dst out
    for batch in (64, 4):
        p = _nt("c", (1 << 20,), batch)
        assert p[0][0] == SRC and p[0][2] & 1, p
        assert p[-1][1] == DST and p[-1][2] & 2, p
