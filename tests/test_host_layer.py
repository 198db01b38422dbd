"""The library's host layer on the CPU (no GPU needed): the C ABI's exports, config.toml
parsing against the reference's own parser, grid / dump errors, the job schedule, the
derived units against the oracle, and the output writers against the reference's own
writers (text image, spectrum and HDF5 byte for byte / dataset for dataset).

The reference-side checks run the reference harnesses built from /root/reference by
oracle/Makefile (oracle/_ref/); they are skipped where that tree is absent.
"""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from phoenix_amd import capi, configs, snapshot as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF64 = os.path.join(ROOT, "oracle", "_ref", "phx_ref_fp64")
REF_OUT = os.path.join(ROOT, "oracle", "_ref", "phx_ref_output_fp64")
H5DIFF = shutil.which("h5diff") or ("/opt/conda/bin/h5diff" if os.path.exists("/opt/conda/bin/h5diff") else None)


# ------------------------------------------------------------------ ABI
def test_every_declared_entry_point_is_exported_and_bound():
    with open(os.path.join(ROOT, "include", "phx.h")) as f:
        text = f.read()
    declared = set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\**\s*(phx_\w+)\s*\(", text, re.M))
    assert len(declared) >= 30, declared
    lib = capi.lib()
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, f"declared in phx.h but not exported by libphx.so: {missing}"
    unbound = sorted(declared - set(capi.SIGNATURES))
    assert not unbound, f"not bound in capi.SIGNATURES: {unbound}"
    assert capi.lib().phx_version().decode()


# ------------------------------------------------------------------ config.toml
def _load_ours(path, overrides=()):
    cfg, model = capi.Config(), capi.Model()
    ov = [o.encode() for o in overrides]
    arr = (ctypes.c_char_p * max(1, len(ov)))(*ov)
    rc = capi.lib().phx_load_config(path.encode(), arr, len(ov), ctypes.byref(cfg), ctypes.byref(model))
    if rc != 0:
        return None, capi.lib().phx_last_error().decode()
    d = {"inclination": model.inclination, "img_width": model.img_width, "img_height": model.img_height,
         "fov_x": model.fov_x, "fov_y": model.fov_y, "r_cam": model.r_cam, "t_init": model.t_init, "pa": model.pa,
         "phi_cam": model.phi_cam, "nfreqs": model.nfreqs, "mass": model.mass, "distance": model.distance,
         "stepsize": model.stepsize, "max_steps": model.max_steps, "horizon_margin": model.horizon_margin,
         "max_order": model.max_order, "m_unit": model.m_unit, "r_low": model.r_low, "r_high": model.r_high,
         "beta_c": model.beta_c, "sigma_cut": model.sigma_cut, "thetae_min": model.thetae_min,
         "thetae_max": model.thetae_max, "rt_outer": model.rt_outer, "nspin": model.nspin, "bpol": model.bpol,
         "nlevels": model.nlevels, "sfc": model.sfc, "outdir": cfg.outdir.decode(), "image": cfg.image.decode(),
         "block_size": cfg.block_size, "dump_file": cfg.dump_file.decode(), "grid_file": cfg.grid_file.decode(),
         "gpu_decode": cfg.gpu_decode, "show_header": cfg.show_header}
    for f in range(model.nfreqs):
        d[f"frequency{f}"] = model.frequencies[f]
    return d, None


def _load_ref(path, overrides=()):
    p = subprocess.run([REF64, path, "--config-only", *overrides], capture_output=True, text=True,
                       env=dict(os.environ, LD_LIBRARY_PATH="/usr/lib/x86_64-linux-gnu"))
    if p.returncode != 0:
        return None, p.stdout.strip()
    d = {}
    for line in p.stdout.splitlines():
        k, v = line.split(" ", 1)
        d[k] = v
    return d, None


CONFIG_CASES = {
    "base": ({}, ()),
    "overrides": ({}, ("camera.inclination=37.5", "plasma.r_high=12", "camera.img_width=64")),
    "freq_list": ({"camera__frequency_list": "[86e9, 230e9, 345e9]"}, ()),
    "freq_linspace": ({"camera__frequency_linspace": "[100e9, 300e9, 5]"}, ()),
    "freq_logspace": ({"camera__frequency_logspace": "[86e9, 690e9, 8]"}, ()),
    "override_freq_form": ({}, ("camera.frequency_list=[86e9, 345e9]",)),
    "pa_phi": ({"camera__pa": "30", "camera__phi": "45"}, ()),
    "text_output": ({"output__image": '"text"', "output__block_size": "10"}, ()),
    # errors
    "unknown_key": ({"camera__bogus": "1"}, ()),
    "two_freq_forms": ({"camera__frequency": "230e9", "camera__frequency_list": "[86e9]"}, None),
    "bad_override": ({}, ("camera.inclination",)),
    "unknown_override": ({}, ("camera.bogus=3",)),
    "dup_freq": ({"camera__frequency_list": "[230e9, 230e9]"}, ()),
    "logspace_n1": ({"camera__frequency_logspace": "[86e9, 690e9, 1]"}, ()),
    "linspace_bad_order": ({"camera__frequency_linspace": "[300e9, 100e9, 3]"}, ()),
    "neg_freq": ({"camera__frequency": "-1"}, ()),
    "bad_image": ({"output__image": '"png"'}, ()),
    "bad_block": ({"output__block_size": "0"}, ()),
    "bad_beta_c": ({"plasma__beta_c": "0"}, ()),
    "not_a_number": ({"camera__inclination": "ten"}, ()),
    "too_many_freqs": ({"camera__frequency_logspace": "[86e9, 690e9, 33]"}, ()),
}


def _write_case(tmp_path, name, over):
    text = configs.config_text("/nonexistent/dump.dat", "/nonexistent/grid.in", **over)
    if name == "two_freq_forms":  # both forms present: config_text would replace one with the other
        text = text.replace("frequency = 230e9", "frequency = 230e9\nfrequency_list = [86e9]")
    p = tmp_path / f"{name}.toml"
    p.write_text(text)
    return str(p)


@pytest.mark.skipif(not os.path.exists(REF64), reason="oracle/_ref/phx_ref_fp64 not built (make -C oracle ref)")
@pytest.mark.parametrize("name", sorted(CONFIG_CASES))
def test_config_matches_reference_parser(tmp_path, name):
    over, ovs = CONFIG_CASES[name]
    path = _write_case(tmp_path, name, over)
    ovs = ovs or ()
    ours, oerr = _load_ours(path, ovs)
    ref, rerr = _load_ref(path, ovs)
    assert (ours is None) == (ref is None), f"ours: {oerr or 'ok'}; reference: {rerr or 'ok'}"
    if ref is None:
        return
    for k, v in ref.items():
        if k in ("outdir", "image", "dump_file", "grid_file"):
            assert ours[k] == v, k
        else:
            assert float(ours[k]) == float(v), (k, ours[k], v)


def test_config_errors_without_reference(tmp_path):
    """error paths whose outcome does not need the reference to state"""
    for name in ("unknown_key", "bad_override", "dup_freq", "bad_image", "bad_block", "bad_beta_c"):
        over, ovs = CONFIG_CASES[name]
        ours, err = _load_ours(_write_case(tmp_path, name, over), ovs or ())
        assert ours is None and err, name
    ours, err = _load_ours(str(tmp_path / "missing.toml"))
    assert ours is None and "missing.toml" in err


# ------------------------------------------------------------------ grid / dump
def test_grid_and_dump_errors(tmp_path, snapdir):
    dump, grid, _ = S.ensure_snapshot("tiny", snapdir)
    cfg = configs.write_config(str(tmp_path / "t.toml"), dump, grid)
    run = capi.Run(cfg)  # loads fine
    assert run.model.nleafs > 0 and run.model.cells > 0
    run.close()
    m = capi.Model()
    with pytest.raises(capi.PhxError) as e:
        capi.check(capi.lib().phx_load_grid(str(tmp_path / "nogrid.in").encode(), ctypes.byref(m)))
    assert e.value.code == -2
    short = tmp_path / "short.dat"
    short.write_bytes(open(dump, "rb").read()[:1000])
    cfg2 = configs.write_config(str(tmp_path / "s.toml"), str(short), grid)
    with pytest.raises(capi.PhxError):
        capi.Run(cfg2)
    # a grid whose base size is not a multiple of the dump's block size
    bad = tmp_path / "badgrid.in"
    bad.write_text(re.sub(r"nxlone1 = (\d+)", lambda mm: f"nxlone1 = {int(mm.group(1)) + 1}", open(grid).read()))
    cfg3 = configs.write_config(str(tmp_path / "b.toml"), dump, str(bad))
    with pytest.raises(capi.PhxError):
        capi.Run(cfg3)


# ------------------------------------------------------------------ schedule, units
def _morton_key(i, j):  # i in the even bits, j in the odd bits (src/host/render/jobs.hpp:33-35)
    k = 0
    for b in range(16):
        k |= ((i >> b) & 1) << (2 * b) | ((j >> b) & 1) << (2 * b + 1)
    return k


@pytest.mark.parametrize("wh", [(16, 16), (12, 20), (64, 8)])
def test_schedule_is_a_morton_ordered_permutation(tmp_path, snapdir, wh):
    dump, grid, _ = S.ensure_snapshot("tiny", snapdir)
    run = capi.Run(configs.write_config(str(tmp_path / "t.toml"), dump, grid))
    m = run.request(img_width=wh[0], img_height=wh[1])
    jobs = run.schedule(True, m)
    assert sorted(jobs.tolist()) == list(range(wh[0] * wh[1]))
    keys = [_morton_key(int(p) // wh[1], int(p) % wh[1]) for p in jobs]
    assert keys == sorted(keys)
    plain = run.schedule(False, m)
    assert plain.tolist() == list(range(wh[0] * wh[1]))
    run.close()


def test_derived_units_match_the_oracle(tmp_path, snapdir, oracle_lib):
    """finalize_model's units and derived scalars (bhac_io.hpp:183-251) against the oracle, which
    is bit-identical to the reference's CPU build"""
    O = oracle_lib
    dump, grid, _ = S.ensure_snapshot("c1", snapdir)
    run = capi.Run(configs.write_config(str(tmp_path / "c.toml"), dump, grid, plasma__m_unit="3e18",
                                        camera__r_cam="2e4", integration__horizon_margin="0.05"))
    o = O.Oracle(dump, grid, 64, m_unit=3e18, r_cam=2e4, horizon_margin=0.05)
    ref = o.info()
    o.close()
    m = run.model
    for k, mine in (("a", m.a), ("hslope", m.hslope), ("L", m.L), ("cut_in", m.cut_in), ("r_out", m.r_out),
                    ("B_unit", m.Bu), ("Ne_unit", m.Ne)):
        assert mine == ref[k], (k, mine, ref[k])
    assert (m.nleafs, m.cells, m.nlevels) == (ref["nleafs"], ref["cells"], ref["nlevels"])
    assert list(m.nx) == ref["nx"] and list(m.ng) == ref["ng"]
    run.close()


# ------------------------------------------------------------------ output writers
def _fake_results(m, seed=3):
    rng = np.random.default_rng(seed)
    n = m.img_width * m.img_height * m.nfreqs
    res = rng.normal(size=(n, 6)) * np.array([1e-3, 1e-4, 1e-4, 1e-5, 1.0, 2.0])
    res[rng.integers(0, n, 5), 0] = np.nan  # non-finite pixels are skipped by the flux sums
    res[rng.integers(0, n, 7)] = 0.0
    return res


@pytest.mark.skipif(not os.path.exists(REF_OUT), reason="oracle/_ref/phx_ref_output_fp64 not built (make -C oracle ref_output)")
@pytest.mark.parametrize("nf", [1, 3])
def test_writers_match_reference_writers(tmp_path, snapdir, nf):
    dump, grid, _ = S.ensure_snapshot("tiny", snapdir)
    freq = "230e9" if nf == 1 else None
    over = dict(camera__img_width="16", camera__img_height="24", output__block_size="8", camera__pa="15")
    if nf == 1:
        over["camera__frequency"] = freq
    else:
        over["camera__frequency_list"] = "[86e9, 230e9, 345e9]"
    cfg = configs.write_config(str(tmp_path / "w.toml"), dump, grid, **over)
    run = capi.Run(cfg)
    res = _fake_results(run.model)
    ours, theirs = tmp_path / "ours", tmp_path / "ref"
    log = run.write_output(res, outdir=str(ours), image="both")
    assert "img_data.h5" in log and "spectrum.dat" in log
    binp = tmp_path / "res.bin"
    res.astype(np.float64).tofile(binp)
    p = subprocess.run([REF_OUT, cfg, str(binp), str(theirs)], capture_output=True, text=True,
                       env=dict(os.environ, LD_LIBRARY_PATH="/usr/lib/x86_64-linux-gnu"))
    assert p.returncode == 0, p.stderr
    names = ["spectrum.dat"] + (["image.dat"] if nf == 1 else [f"image_{f}.dat" for f in range(nf)])
    for name in names:
        assert (ours / name).read_bytes() == (theirs / name).read_bytes(), name
    if H5DIFF:
        d = subprocess.run([H5DIFF, str(ours / "img_data.h5"), str(theirs / "img_data.h5")], capture_output=True,
                           text=True, env=dict(os.environ, LD_LIBRARY_PATH="/opt/conda/lib"))
        assert d.returncode == 0, d.stdout + d.stderr
    run.close()


def test_writer_falls_back_to_text_when_hdf5_layout_is_invalid(tmp_path, snapdir):
    dump, grid, _ = S.ensure_snapshot("tiny", snapdir)
    run = capi.Run(configs.write_config(str(tmp_path / "w.toml"), dump, grid, camera__img_width="10",
                                        camera__img_height="10", output__block_size="3"))
    res = _fake_results(run.model)
    log = run.write_output(res, outdir=str(tmp_path / "o"), image="hdf5")
    assert "Failed to write" in log and (tmp_path / "o" / "image.dat").exists()
    rows = np.loadtxt(tmp_path / "o" / "image.dat")
    assert rows.shape == (100, 8)
    run.close()


# ------------------------------------------------------------------ device tables
def _float_table(text, name):
    body = re.search(name + r"\[3\]\[24\][^=]*=\s*\{(.*?)\};", text, re.S).group(1)
    rows = re.findall(r"\{([^{}]*)\}", body)
    return np.array([[float(v) for v in r.split(",") if v.strip()] for r in rows], dtype=np.float32)


def test_bessel_tables_padded_with_zeros_give_the_same_clenshaw_sums():
    """The FP32 march evaluates the Bessel-ratio Chebyshev fits over all 24 table entries of a piece
    instead of the fit's own length: the entries past the length are 0, the recurrence stays exactly
    at 0 through them from its zero start, so the sums are bit-identical (float32 arithmetic here)."""
    with open(os.path.join(ROOT, "phoenix_amd", "csrc", "phx_bessel_table.h")) as f:
        text = f.read()
    nterm = [int(v) for v in re.search(r"PHX_BES_NTERM\[3\] = \{([^}]*)\}", text).group(1).split(",")]
    z = np.linspace(-1.0, 1.0, 4001, dtype=np.float32)
    two = np.float32(2)
    for name in ("PHX_BES_K1K2_F", "PHX_BES_K0K2_F"):
        tab = _float_table(text, name)
        assert tab.shape == (3, 24)
        for p in range(3):
            assert np.all(tab[p, nterm[p]:] == 0), (name, p)

            def clenshaw(top):
                b1 = np.zeros_like(z)
                b2 = np.zeros_like(z)
                for j in range(top, 0, -1):
                    t = two * z * b1 - b2 + tab[p, j]
                    b2, b1 = b1, t
                return tab[p, 0] + z * b1 - b2

            short, padded = clenshaw(nterm[p] - 1), clenshaw(23)
            assert np.array_equal(short.view(np.uint32), padded.view(np.uint32)), (name, p)
