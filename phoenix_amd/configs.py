"""config.toml files in the reference's key set (``input/config.toml``,
``src/host/io/config.hpp:301-370``) for the named workloads."""
from __future__ import annotations

# The base workload: 230 GHz, inclination 10 deg, m_unit 1.2e18, R_high 3, camera at
# r = 1e4, fov 40 x 40 Rg, stepsize 0.01, max_steps 1e4 (the reference's defaults
# elsewhere).
BASE = {
    "visual": {"show_header": "false"},
    "output": {"dir": '"output"', "image": '"none"', "block_size": "64"},
    "camera": {"inclination": "10", "img_width": "128", "img_height": "128", "fov_x": "40",
               "fov_y": "40", "r_cam": "1e4", "t_init": "0", "frequency": "230e9"},
    "black_hole": {"mass": "4.14e6", "distance": "8.127"},
    "integration": {"stepsize": "0.01", "max_steps": "1e4", "horizon_margin": "0.01",
                    "max_order": "100"},
    "plasma": {"m_unit": "1.2e18", "r_low": "1", "r_high": "3", "sigma_cut": "1.0",
               "thetae_min": "1e-3", "thetae_max": "100.0", "rt_outer": "50.0"},
    "grmhd": {"dump_file": '""', "grid_file": '""', "gpu_decode": "false", "nspin": "3",
              "bpol": "1", "sfc": "true", "nlevels": "-1"},
}


def config_text(dump: str, grid: str, **over) -> str:
    """TOML text; keyword overrides are ``section__key=value`` (value as TOML text).
    Any ``camera__frequency*`` override replaces the single ``frequency`` key."""
    cfg = {s: dict(kv) for s, kv in BASE.items()}
    cfg["grmhd"]["dump_file"] = f'"{dump}"'
    cfg["grmhd"]["grid_file"] = f'"{grid}"'
    for k, v in over.items():
        sec, key = k.split("__", 1)
        if sec == "camera" and key.startswith("frequenc"):
            for fk in [f for f in cfg["camera"] if f.startswith("frequenc")]:
                del cfg["camera"][fk]
        cfg.setdefault(sec, {})[key] = str(v)
    lines = []
    for sec, kv in cfg.items():
        lines.append(f"[{sec}]")
        for k, v in kv.items():
            lines.append(f"{k} = {v}")
        lines.append("")
    return "\n".join(lines)


def write_config(path: str, dump: str, grid: str, **over) -> str:
    with open(path, "w") as f:
        f.write(config_text(dump, grid, **over))
    return path
