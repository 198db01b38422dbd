"""The reference's own driver, src/main.cpp unchanged, built with device::render from
phoenix_amd/integration/kernel_phx.cpp on libphx (oracle/Makefile, target ref_main_phx).

CPU: the driver links and resolves every libphx symbol it binds.  GPU: it renders a snapshot
end to end (the reference's config loader, grid and dump loader, Morton schedule and output
writer around libphx's device path), and its text image and spectrum equal libphx's own
driver's at the same precision: the reference's loader fills the same tables as phx_load_dump,
and kernel_phx.cpp hands them over unchanged."""
import os
import subprocess

import numpy as np
import pytest

from phoenix_amd import configs, snapshot as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MAIN = {acc: os.path.join(ROOT, "oracle", "_ref", f"phoenix_main_phx_{acc.lower()}") for acc in ("FP32", "FP64")}
CLI = os.path.join(ROOT, "phoenix_amd", "phoenix")
BOUND = ("phx_dev_warmup", "phx_dev_render", "phx_model_derive", "phx_last_error")


def _need(path):
    if not os.path.exists(path):
        pytest.skip(f"{os.path.relpath(path, ROOT)} is built from the reference sources (oracle/Makefile ref_main_phx)")


def _run(cmd, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300)
    print(p.stderr[-2000:])
    return p


@pytest.mark.parametrize("acc", ["FP32", "FP64"])
def test_reference_driver_links_libphx(acc):
    exe = REF_MAIN[acc]
    _need(exe)
    # every symbol resolved at start-up (LD_BIND_NOW), then the usage path, which touches no device
    p = _run([exe, "--help"], {"LD_BIND_NOW": "1"})
    assert p.returncode == 0 and "usage" in p.stderr + p.stdout
    und = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True, check=True).stdout
    exported = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "phoenix_amd", "libphx.so")],
                              capture_output=True, text=True, check=True).stdout.split()
    for sym in BOUND:
        assert sym in und.split(), sym
        assert sym in exported, sym


def _config(tmp_path, snapdir, out, gpu_decode, nf):
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    freqs = "[230e9, 345e9]" if nf == 2 else "[230e9]"
    return configs.write_config(str(tmp_path / "c.toml"), dump, grid, output__dir=f'"{out}"', output__image='"text"',
                                camera__img_width="40", camera__img_height="28", camera__frequency_list=freqs,
                                grmhd__gpu_decode="true" if gpu_decode else "false")


def _outputs(d, nf):
    names = ["spectrum.dat"] + (["image.dat"] if nf == 1 else [f"image_{f}.dat" for f in range(nf)])
    return {n: (d / n).read_text() for n in names}


@pytest.mark.gpu
@pytest.mark.parametrize("acc,precision", [("FP32", "FP32"), ("FP64", "FP64"), ("FP64", "MIXED")])
def test_reference_driver_on_libphx_equals_libphx_driver(snapdir, tmp_path, acc, precision):
    exe = REF_MAIN[acc]
    _need(exe)
    _need(CLI)
    nf = 2
    out_ref, out_cli = tmp_path / "ref", tmp_path / "cli"
    cfg = _config(tmp_path, snapdir, out_ref, True, nf)
    print(f"reference driver {acc} build, libphx precision {precision}", flush=True)
    p = _run([exe, cfg, "camera.inclination=35"], {"PHX_ACCURACY": precision})
    assert p.returncode == 0
    assert "Flux at 2.3000e+11 Hz" in p.stderr + p.stdout
    q = _run([CLI, cfg, "camera.inclination=35", f"output.dir={out_cli}"], {"PHX_ACCURACY": precision})
    assert q.returncode == 0
    a, b = _outputs(out_ref, nf), _outputs(out_cli, nf)
    for name in a:
        assert a[name] == b[name], name
    spec = np.loadtxt(out_ref / "spectrum.dat", ndmin=2)
    assert spec.shape == (nf, 7) and np.all(np.isfinite(spec)) and np.all(spec[:, 1] > 0)


@pytest.mark.gpu
def test_reference_driver_host_decoded_primitives(snapdir, tmp_path):
    """grmhd.gpu_decode = false: the reference decodes the dump on the host and kernel_phx.cpp
    passes its primitives to libphx; the spectrum agrees with the device-decoded render."""
    exe = REF_MAIN["FP64"]
    _need(exe)
    nf = 1
    outs = {}
    for gd in (False, True):
        out = tmp_path / f"gd{int(gd)}"
        cfg = _config(tmp_path, snapdir, out, gd, nf)
        p = _run([exe, cfg], {"PHX_ACCURACY": "FP64"})
        assert p.returncode == 0
        outs[gd] = np.loadtxt(out / "spectrum.dat", ndmin=2)
    # host and device decode differ in the last bits of a few primitives (pow, sqrt of the
    # conserved-to-primitive inversion), so the fluxes agree to rounding, not bit for bit
    np.testing.assert_allclose(outs[False][:, 1:5], outs[True][:, 1:5], rtol=1e-6,
                               atol=1e-9 * abs(outs[True][0, 1]))
