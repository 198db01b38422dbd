"""The captured render (phx_graph_*) through the C ABI: the allocation-free render's launch sequence
recorded once into a HIP graph and replayed, for the same parameter point and for new sweep points
of the same camera (m_unit, R_high, frequency values).  Every replay must equal the allocating
render of the same request bit for bit (the same kernels march the same rays), with the program's
own eager renders, and an allocation-free render that shares the graph's device array, between
the replays; requests that change the camera, the image or the number of frequencies are refused."""
import numpy as np
import pytest

from phoenix_amd import capi, configs, snapshot as S
from tests.devbuf import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(snapdir, tmp_path_factory):
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    run = capi.Run(configs.write_config(str(tmp_path_factory.mktemp("gr") / "c1.toml"), dump, grid))
    snap = capi.Snapshot(run, capi.TABLE_FP32 | capi.TABLE_FP64)
    yield run, snap
    snap.close()
    run.close()


@pytest.mark.parametrize("precision,freqs", [(capi.MIXED, (230e9,)), (capi.FP32, (230e9, 345e9)),
                                             (capi.FP64, (86e9,))])
def test_graph_replays_equal_eager_renders(c1, precision, freqs):
    run, snap = c1
    w, h, incl = 48, 40, 30.0
    cam = dict(img_width=w, img_height=h, inclination=incl)
    points = [run.request(**cam, freqs=freqs),
              run.request(**cam, freqs=tuple(1.5 * f for f in freqs), m_unit=3e18, r_high=20.0),
              run.request(**cam, freqs=tuple(0.5 * f for f in freqs), m_unit=6e17, r_high=1.5)]
    n = w * h * len(freqs)
    buf, other = DeviceBuffer(n), DeviceBuffer(24 * 24)
    g = snap.graph(points[0], precision, buf.ptr.value)
    try:
        for rep in range(2):
            for k, rq in enumerate(points):
                buf.clear()
                info = g.launch(None if rep == 0 and k == 0 else rq)
                got = buf.get().reshape(w * h, len(freqs), 6)
                # eager work between the replays: an allocating render of the same point, and an
                # allocation-free render of another request in the graph's device array
                ref, iref = snap.render(rq, precision)
                snap.render_noalloc(run.request(img_width=24, img_height=24, freqs=(345e9,)), precision,
                                    other.ptr.value)
                print(f"\nfp{precision} replay {rep}.{k}: kernel {info['kernel_s'] * 1e3:.2f} ms, "
                      f"steps {info['steps']} (eager {iref['steps']})")
                assert np.array_equal(got, ref), f"replay {rep}.{k} differs from the eager render"
                assert (info["steps"], info["nplasma"], info["bad"]) == (iref["steps"], iref["nplasma"], iref["bad"])
                assert info["vgprs"] > 0 and info["blocks"] > 0
        # the same point again, without new arguments: the last point's result
        buf.clear()
        g.launch()
        ref, _ = snap.render(points[-1], precision)
        assert np.array_equal(buf.get().reshape(w * h, len(freqs), 6), ref)
    finally:
        g.close()
        buf.free()
        other.free()


def test_graph_refuses_another_camera_or_image(c1):
    run, snap = c1
    rq = run.request(img_width=32, img_height=32, inclination=20.0)
    buf = DeviceBuffer(32 * 32 * 2)
    g = snap.graph(rq, capi.FP32, buf.ptr.value)
    try:
        for bad in (run.request(img_width=32, img_height=16, inclination=20.0),
                    run.request(img_width=32, img_height=32, inclination=25.0),
                    run.request(img_width=32, img_height=32, inclination=20.0, freqs=(86e9, 230e9)),
                    run.request(img_width=32, img_height=32, inclination=20.0, stepsize=0.02)):
            with pytest.raises(capi.PhxError) as e:
                g.launch(bad)
            assert e.value.code == -1
        g.launch(run.request(img_width=32, img_height=32, inclination=20.0, m_unit=2e18))  # a sweep point
    finally:
        g.close()
        buf.free()
    with pytest.raises(capi.PhxError) as e:  # beyond the allocation-free render's job limit
        capi.Graph(snap, run.request(img_width=1100, img_height=1000), capi.FP32, 16)
    assert e.value.code == -4
