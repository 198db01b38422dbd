"""The allocation-free render (phx_snapshot_render_noalloc) through the C ABI: into a device buffer
the caller owns, with the library's fixed device array for its per-render buffers.  Its results
must equal the allocating entry point's (phx_snapshot_render) bit for bit (the same kernels march
the same rays), on two calls in a row, the second with a smaller image, in every precision; sizes
beyond its limits are refused with PHX_E_RANGE."""
import numpy as np
import pytest

from phoenix_amd import capi, configs, snapshot as S
from tests.devbuf import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(snapdir, tmp_path_factory):
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    run = capi.Run(configs.write_config(str(tmp_path_factory.mktemp("na") / "c1.toml"), dump, grid))
    snap = capi.Snapshot(run, capi.TABLE_FP32 | capi.TABLE_FP64)
    yield run, snap
    snap.close()
    run.close()


@pytest.mark.parametrize("precision,freqs", [(capi.MIXED, (230e9,)), (capi.FP32, (230e9, 345e9)),
                                             (capi.FP64, (86e9,))])
def test_noalloc_equals_allocating_render(c1, precision, freqs):
    run, snap = c1
    for w, h, incl in ((64, 64, 10.0), (40, 24, 60.0)):  # the second call is smaller (and not a square)
        rq = run.request(img_width=w, img_height=h, inclination=incl, freqs=freqs)
        n = w * h * len(freqs)
        buf = DeviceBuffer(n)
        info = snap.render_noalloc(rq, precision, buf.ptr.value)
        got = buf.get().reshape(w * h, len(freqs), 6)
        buf.free()
        ref, iref = snap.render(rq, precision)
        print(f"\nfp{precision} {w}x{h} nf={len(freqs)}: kernel {info['kernel_s'] * 1e3:.2f} ms, "
              f"steps {info['steps']} (allocating render {iref['steps']})")
        assert np.array_equal(got, ref)
        assert (info["steps"], info["nplasma"], info["bad"]) == (iref["steps"], iref["nplasma"], iref["bad"])


def test_noalloc_limits(c1):
    run, snap = c1
    buf = DeviceBuffer(8 * 8 * 4)
    with pytest.raises(capi.PhxError) as e:  # 1100 x 1000 jobs > 1,048,576
        snap.render_noalloc(run.request(img_width=1100, img_height=1000), capi.FP32, buf.ptr.value)
    assert e.value.code == -4
    with pytest.raises(capi.PhxError) as e:  # four frequencies per lane: a kernel with per-lane scratch
        snap.render_noalloc(run.request(img_width=8, img_height=8, freqs=(86e9, 230e9, 345e9, 690e9)),
                            capi.MIXED, buf.ptr.value)
    assert e.value.code == -4
    buf.free()
    with pytest.raises(capi.PhxError):
        snap.render_noalloc(run.request(img_width=8, img_height=8), capi.FP32, 0)  # no output buffer
