"""The command-line driver (phoenix_amd/phoenix, the counterpart of the reference's main.cpp)
end to end on the GPU: config + overrides -> render -> HDF5 / text image and spectrum, and the
--batch form; its outputs against the same render through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

from phoenix_amd import capi, configs, snapshot as S

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(capi.HERE), "phoenix_amd", "phoenix")


def _run_cli(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    p = subprocess.run([CLI, *args], capture_output=True, text=True, env=env, timeout=300)
    print(p.stderr[-3000:])
    return p


@pytest.mark.skipif(not os.path.exists(CLI), reason="phoenix driver not built")
@pytest.mark.parametrize("accuracy,precision", [("FP32", capi.FP32), ("MIXED", capi.MIXED)])
def test_cli_single_image_matches_the_abi(snapdir, tmp_path, accuracy, precision):
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    out = tmp_path / "out"
    cfg = configs.write_config(str(tmp_path / "c.toml"), dump, grid, output__dir=f'"{out}"', output__image='"both"',
                               output__block_size="8", camera__img_width="32", camera__img_height="24",
                               camera__frequency_list="[230e9, 345e9]")
    p = _run_cli([cfg, "camera.inclination=30"], {"PHX_ACCURACY": accuracy})
    assert p.returncode == 0
    for name in ("img_data.h5", "image_0.dat", "image_1.dat", "spectrum.dat"):
        assert (out / name).exists(), name
    assert "Flux at 2.3000e+11 Hz" in p.stderr
    # the same render through the C ABI, written by the same writer
    run = capi.Run(cfg, overrides=["camera.inclination=30"])
    res, _ = run.dev_render(precision)
    ref = tmp_path / "abi"
    run.write_output(res, outdir=str(ref), image="text")
    assert (out / "spectrum.dat").read_text() == (ref / "spectrum.dat").read_text()
    assert (out / "image_1.dat").read_text() == (ref / "image_1.dat").read_text()
    run.close()


@pytest.mark.skipif(not os.path.exists(CLI), reason="phoenix driver not built")
def test_cli_batch(snapdir, tmp_path):
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    out = tmp_path / "out"
    cfg = configs.write_config(str(tmp_path / "c.toml"), dump, grid, output__dir=f'"{out}"', output__image='"text"',
                               camera__img_width="16", camera__img_height="16")
    batch = tmp_path / "batch.txt"
    batch.write_text("camera.inclination=20 plasma.r_high=10\n# a comment line\n\n"
                     "camera.inclination=70 camera.img_width=24 camera.frequency_list=[86e9,230e9]\n")
    p = _run_cli([cfg, "--batch", str(batch)])
    assert p.returncode == 0
    run = capi.Run(cfg)
    snap = capi.Snapshot(run, capi.TABLE_FP32)
    for k, (over, nf) in enumerate([(dict(inclination=20.0, r_high=10.0), 1),
                                    (dict(inclination=70.0, img_width=24, freqs=(86e9, 230e9)), 2)]):
        rq = run.request(**over)
        alone, _ = snap.render(rq, capi.FP32)
        S_ref = run.flux(alone, rq)
        spec = np.loadtxt(out / str(k) / "spectrum.dat", ndmin=2)
        assert spec.shape[0] == nf
        np.testing.assert_allclose(spec[:, 1:5], S_ref, rtol=1e-6, atol=1e-12 * np.abs(S_ref[:, :1]).max())
    snap.close()
    run.close()


@pytest.mark.skipif(not os.path.exists(CLI), reason="phoenix driver not built")
def test_cli_reports_errors(tmp_path):
    p = _run_cli([str(tmp_path / "missing.toml")])
    assert p.returncode != 0 and "FAIL" in p.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="phoenix driver not built")
def test_cli_serve(snapdir, tmp_path):
    """--serve FILE: requests with arrival times, on two resident snapshots (a line names the other
    dump; override values are taken as written, without quotes, as the reference's config.hpp:246-262
    takes them), through the serving loop; every request's spectrum against the same request
    rendered alone through the C ABI."""
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    d2, g2, _ = S.ensure_snapshot("c1_a05", snapdir)
    out = tmp_path / "out"
    cfg = configs.write_config(str(tmp_path / "c.toml"), dump, grid, output__dir=f'"{out}"', output__image='"none"',
                               camera__img_width="16", camera__img_height="16")
    serve = tmp_path / "serve.txt"
    serve.write_text(f"0.0 camera.inclination=20\n"
                     f"0.05 camera.inclination=70 camera.img_width=24 camera.frequency_list=[86e9,230e9]\n"
                     f"0.02 grmhd.dump_file={d2} grmhd.grid_file={g2} camera.inclination=45 plasma.r_high=20\n")
    p = _run_cli([cfg, "--serve", str(serve)], {"PHX_SERVE_QUANTUM": "100"})
    assert p.returncode == 0
    assert "Served 3 requests" in p.stderr
    cases = [(cfg, dict(inclination=20.0)), (cfg, dict(inclination=70.0, img_width=24, freqs=(86e9, 230e9))),
             (configs.write_config(str(tmp_path / "c2.toml"), d2, g2, camera__img_width="16", camera__img_height="16"),
              dict(inclination=45.0, r_high=20.0))]
    for k, (c, over) in enumerate(cases):
        run = capi.Run(c)
        snap = capi.Snapshot(run, capi.TABLE_FP32)
        rq = run.request(**over)
        alone, _ = snap.render(rq, capi.FP32)
        S_ref = run.flux(alone, rq)
        spec = np.loadtxt(out / str(k) / "spectrum.dat", ndmin=2)
        np.testing.assert_allclose(spec[:, 1:5], S_ref, rtol=1e-6, atol=1e-12 * np.abs(S_ref[:, :1]).max())
        snap.close()
        run.close()
