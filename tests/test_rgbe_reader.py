"""EnvironmentLight(path), the Radiance RGBE (.hdr) reader of the C++ host
interface, through `render_demo --dump-hdr` (no GPU).  It stands in for the
reference's stb_image load of hrdi/*.hdr (load_texture_float,
CRT/dTexture.cu:189-267), which flips rows (bottom row first).

Files are written here by an encoder that follows the format (flat
scanlines; new-style run-length scanlines with runs and dumps), with known
texels; each decoded texel must equal mantissa * 2^(exponent - 136) exactly
(zero for a zero exponent), rows bottom first.  Truncated files, runs or
dumps past the end of a scanline, zero-length dumps and bad headers are
rejected.  When the reference tree is present, its own hrdi/*.hdr files are
read and their sizes checked against their headers."""
import glob
import os
import subprocess

import numpy as np
import pytest

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "firefly_amd", "render_demo")
REF_HDR = sorted(glob.glob("/root/reference/hrdi/*.hdr"))


def header(w, h, magic=b"#?RADIANCE", fmt=b"FORMAT=32-bit_rle_rgbe", res=None):
    out = magic + b"\n# written by the test\n" + fmt + b"\nEXPOSURE=1.0\n\n"
    return out + (res if res is not None else b"-Y %d +X %d\n" % (h, w))


def flat(px):
    """px [h][w][4] uint8, top scanline first."""
    return px.astype(np.uint8).tobytes()


def rle_plane(vals, runs_at):
    """One byte plane of a scanline as runs (starting where runs_at says) and dumps."""
    out = bytearray()
    x, w = 0, len(vals)
    while x < w:
        if x in runs_at:
            n = 1
            while x + n < w and n < 127 and vals[x + n] == vals[x]:
                n += 1
            if n >= 2:
                out += bytes([128 + n, vals[x]])
                x += n
                continue
        n = 1
        while x + n < w and n < 128 and (x + n) not in runs_at:
            n += 1
        out += bytes([n]) + bytes(vals[x:x + n])
        x += n
    return bytes(out)


def rle(px, seed=0):
    h, w, _ = px.shape
    rng = np.random.default_rng(seed)
    out = bytearray()
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for ch in range(4):
            vals = [int(v) for v in px[y, :, ch]]
            # runs wherever the plane repeats, plus some places a run could start
            runs_at = {x for x in range(w) if x + 1 < w and vals[x] == vals[x + 1]} | \
                {int(x) for x in rng.integers(0, w, 3)}
            out += rle_plane(vals, runs_at)
    return bytes(out)


def expected(px):
    """Texels [h][w][3] float32, bottom row first."""
    m = px[..., :3].astype(np.float64)
    e = px[..., 3].astype(np.int64)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (m * scale[..., None]).astype(np.float32)[::-1]


def load(tmp_path, data, name="env.hdr"):
    path = tmp_path / name
    path.write_bytes(data)
    out = tmp_path / (name + ".bin")
    r = subprocess.run([DEMO, "--dump-hdr", str(path), str(out)], capture_output=True, text=True, timeout=60)
    if r.returncode != 0:
        return None, r
    raw = out.read_bytes()
    w, h = np.frombuffer(raw, np.int32, 2)
    return np.frombuffer(raw, np.float32, offset=8).reshape(h, w, 3), r


def image(w, h, seed, runs=True):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    px[..., 3] = rng.integers(100, 160, (h, w))
    px[rng.random((h, w)) < 0.05, 3] = 0       # zero exponent: black texel
    px[0, 0, 3], px[0, 1, 3] = 255, 1          # extreme exponents (1 gives a subnormal float)
    if runs:  # long constant stretches so the encoder writes runs
        px[:, 5:40] = px[:, 5:6]
        px[h // 2, :, 3] = 128
    px[:, 0, 0] = np.where(px[:, 0, 0] == 2, 3, px[:, 0, 0])  # a flat scanline never starts 2, 2
    return px


@pytest.fixture(scope="module", autouse=True)
def demo():
    assert os.path.exists(DEMO), "render_demo not built"


@pytest.mark.parametrize("w,h", [(64, 32), (300, 7), (8, 3)])
def test_rle_scanlines_decode_exactly(tmp_path, w, h):
    px = image(w, h, seed=w + h)
    got, r = load(tmp_path, header(w, h) + rle(px, seed=w))
    assert got is not None, r.stderr
    assert got.shape == (h, w, 3)
    assert np.array_equal(got.view(np.uint32), expected(px).view(np.uint32))


@pytest.mark.parametrize("w,h", [(64, 32), (5, 4), (40000, 2)])
def test_flat_scanlines_decode_exactly(tmp_path, w, h):
    """Flat scanlines: any width below 8 or from 32768 up, and files whose
    scanlines do not start with the run-length marker."""
    px = image(w, h, seed=3 * w + h, runs=False)
    got, r = load(tmp_path, header(w, h) + flat(px))
    assert got is not None, r.stderr
    assert np.array_equal(got.view(np.uint32), expected(px).view(np.uint32))


def test_rows_are_stored_bottom_first(tmp_path):
    w, h = 16, 4
    px = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        px[y, :, :3] = 10 * (y + 1)
        px[y, :, 3] = 136  # scale 1
    got, r = load(tmp_path, header(w, h) + rle(px))
    assert got is not None, r.stderr
    assert [float(got[row, 0, 0]) for row in range(h)] == [40.0, 30.0, 20.0, 10.0]


def test_rgbe_magic_alternative(tmp_path):
    px = image(16, 2, seed=9)
    got, r = load(tmp_path, header(16, 2, magic=b"#?RGBE") + rle(px))
    assert got is not None, r.stderr


def _cases():
    w, h = 32, 6
    px = image(w, h, seed=11)
    good = rle(px)
    body_start = len(header(w, h))
    cases = {
        "truncated_in_rle_data": header(w, h) + good[: len(good) // 2],
        "truncated_last_byte": header(w, h) + good[:-1],
        "truncated_after_run_count": None,
        "truncated_in_dump": None,
        "truncated_flat": header(w, h) + flat(px)[:-5],
        "empty_after_header": header(w, h),
        "run_past_end": header(w, h) + bytes([2, 2, 0, w, 128 + w + 1, 7]),
        "dump_past_end": header(w, h) + bytes([2, 2, 0, w, w + 1]) + bytes(w + 1),
        "zero_length_dump": header(w, h) + bytes([2, 2, 0, w, 0]),
        "no_magic": header(w, h, magic=b"#?JPEG") + good,
        "no_format": header(w, h, fmt=b"FORMAT=32-bit_rle_xyze") + good,
        "flipped_x_resolution": header(w, h, res=b"-Y %d -X %d\n" % (h, w)) + good,
        "y_up_resolution": header(w, h, res=b"+Y %d +X %d\n" % (h, w)) + good,
        "zero_width": header(w, h, res=b"-Y 6 +X 0\n") + good,
        "huge_image": header(w, h, res=b"-Y 65536 +X 65536\n") + good,
        "trailing_garbage_on_resolution": header(w, h, res=b"-Y 6 +X 32 x\n") + good,
    }
    # a run whose value byte is missing; a dump cut short
    cases["truncated_after_run_count"] = header(w, h) + bytes([2, 2, 0, w, 128 + 4])
    cases["truncated_in_dump"] = header(w, h) + bytes([2, 2, 0, w, 10, 1, 2, 3])
    assert body_start > 0
    return cases


@pytest.mark.parametrize("name", sorted(_cases()))
def test_malformed_files_are_rejected(tmp_path, name):
    got, r = load(tmp_path, _cases()[name])
    assert got is None and r.returncode == 1 and "EnvironmentLight" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.skipif(not REF_HDR, reason="the reference's hrdi/ directory is not present")
@pytest.mark.parametrize("path", REF_HDR, ids=[os.path.basename(p) for p in REF_HDR])
def test_reference_hdr_files(tmp_path, path):
    """The reference's own environment maps load with the size their
    resolution line states; texels are finite and non-negative and the map
    has a sampling distribution (not all zero)."""
    with open(path, "rb") as f:
        head = f.read(4096)
    res = [ln for ln in head.split(b"\n") if ln.startswith(b"-Y ")][0].split()
    h, w = int(res[1]), int(res[3])
    got, r = load(tmp_path, open(path, "rb").read(), name=os.path.basename(path))
    assert got is not None, r.stderr
    assert got.shape == (h, w, 3)
    assert np.isfinite(got).all() and (got >= 0).all() and got.sum() > 0
