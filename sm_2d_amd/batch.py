"""Batch sweep driver: `python -m sm_2d_amd.batch -c batch_config.yaml [--dry-run]`.

The reference's batch_run.py (:26-280, batch_config.yaml) expands a template INI over a
Cartesian product of parameter values and starts one `run_simulation` process per
combination, one after another.  This driver reads the same YAML schema and writes the same
per-run directories (`<output_base>/batch_<name>/` with `sim_run.ini`, `dose_2d.txt`,
`pdd.txt`), but transports every group of runs that share a grid and transport settings in
one sm2d_run_batch launch on one context: the spots of a group move through each iteration
together.  Each run's files are byte-identical to what `sm2d_run <its sim_run.ini>` writes,
because a batched spot's FP64 EdepC equals its single-spot EdepC and the same writers are
used.

Keys of the YAML file: `template` (INI path, relative to the working directory as in the
reference), `output_base`, `sweep` (groups of `section` plus lists of values), `execution`
(`parallel` and `timeout` are accepted and not needed: there is one process), `plot`
(not implemented here; plotting is outside the transport step).
"""
from __future__ import annotations

import argparse
import copy
import os
import sys
import time
from itertools import product

import yaml

from . import io as sio

MAX_SPOTS_PER_LAUNCH = 256


class TemplateIni:
    """The template INI as batch_run.py:26-71 models it: '#' comment lines and lines outside a
    section are dropped, sections keep their order, a modification adds or replaces a key,
    and saving writes `[section]`, `key = value` lines and a blank line per section."""

    def __init__(self, path: str):
        self.sections: dict[str, dict[str, str]] = {}
        cur = None
        with open(path, encoding="utf-8") as f:
            for line in f:
                line = line.strip()
                if not line or line.startswith("#"):
                    continue
                if line.startswith("[") and line.endswith("]"):
                    cur = line
                    self.sections[cur] = {}
                elif "=" in line and cur:
                    k, v = line.split("=", 1)
                    self.sections[cur][k.strip()] = v.strip()

    def modify(self, section: str, key: str, value) -> None:
        self.sections.setdefault(section, {})[key] = str(value)

    def save(self, path: str) -> None:
        with open(path, "w", encoding="utf-8") as f:
            for sec, kv in self.sections.items():
                f.write(f"{sec}\n")
                for k, v in kv.items():
                    f.write(f"{k} = {v}\n")
                f.write("\n")


def generate_combinations(sweep: dict) -> list[dict]:
    """ParameterSweep.generate_combinations (batch_run.py:77-137): the product of the groups'
    value lists (within a group and across groups, in the file's order), named by joining
    `<key without _mm/_MeV><value>` parts with '_'."""
    groups = []
    for _, gc in (sweep or {}).items():
        section = gc.get("section", "")
        params = {k: v for k, v in gc.items() if k != "section"}
        if not params:
            continue
        keys = list(params)
        values = [params[k] if isinstance(params[k], list) else [params[k]] for k in keys]
        groups.append([{"section": section, "params": dict(zip(keys, combo))}
                       for combo in product(*values)])
    out = []
    if groups:
        for multi in product(*groups):
            parts = [f"{k.replace('_mm', '').replace('_MeV', '')}{v}"
                     for mod in multi for k, v in mod["params"].items()]
            out.append({"name": "_".join(parts), "modifications": list(multi)})
    return out


class BatchRunner:
    """Prepares the runs of a batch configuration and executes them in batched launches."""

    def __init__(self, config_file: str = "batch_config.yaml", lib=None, device: int = 0):
        with open(config_file, encoding="utf-8") as f:
            self.batch_config = yaml.safe_load(f) or {}
        self.template = TemplateIni(self.batch_config["template"])
        self.lib = lib
        self.device = device
        self.runs: list[dict] = []
        self.completed: list[dict] = []
        self.failed: list[dict] = []

    def prepare_runs(self) -> list[dict]:
        base = self.batch_config.get("output_base", "results/batch")
        runs = []
        for idx, combo in enumerate(generate_combinations(self.batch_config.get("sweep", {}))):
            name = f"batch_{combo['name']}"
            run_dir = os.path.join(base, name)
            cfg = copy.deepcopy(self.template)
            for mod in combo["modifications"]:
                for k, v in mod["params"].items():
                    cfg.modify(mod["section"], k, v)
            cfg.modify("[output]", "output_dir", os.path.abspath(run_dir))
            runs.append({"index": idx, "name": name, "dir": run_dir, "config": cfg, "params": combo})
        self.runs = runs
        return runs

    def _load(self, run: dict):
        os.makedirs(run["dir"], exist_ok=True)
        path = os.path.join(run["dir"], "sim_run.ini")
        run["config"].save(path)
        cfg = sio.config_load(self.lib, path)
        sio.config_validate(self.lib, cfg)
        g, tp, sp = sio.config_runtime(self.lib, cfg)
        return cfg, g, tp, sp

    def run_all(self, max_spots: int = MAX_SPOTS_PER_LAUNCH) -> dict:
        """Runs every prepared run; returns timing and grouping figures."""
        if self.lib is None:
            from .runtime import load_library
            self.lib = load_library()
        if not self.runs:
            self.prepare_runs()
        groups: dict[tuple, list] = {}
        for run in self.runs:
            try:
                cfg, g, tp, sp = self._load(run)
            except Exception as e:  # an invalid run fails alone, as a failed process would
                print(f"[FAIL] {run['name']}: {e}")
                self.failed.append(run)
                continue
            key = (bytes(g), bytes(tp), tuple(cfg.groups()))
            groups.setdefault(key, []).append((run, cfg, g, tp, sp))
        t0 = time.perf_counter()
        launches = 0
        for members in groups.values():
            for i in range(0, len(members), max_spots):
                chunk = members[i:i + max_spots]
                launches += 1
                try:
                    self._run_group(chunk)
                    for run, *_ in chunk:
                        print(f"[OK] {run['name']}")
                        self.completed.append(run)
                except Exception as e:
                    for run, *_ in chunk:
                        print(f"[FAIL] {run['name']}: {e}")
                        self.failed.append(run)
        secs = time.perf_counter() - t0
        print(f"\ncompleted: {len(self.completed)}/{len(self.runs)} in {launches} batched launch(es), "
              f"{secs:.2f} s")
        for run in self.failed:
            print(f"  failed: {run['name']}")
        return {"runs": len(self.runs), "completed": len(self.completed), "launches": launches,
                "seconds": secs}

    def _run_group(self, chunk: list) -> None:
        from .engine import Engine
        _, cfg0, g, tp, _ = chunk[0]
        tab = self.lib.make_tables(g.N_theta, g.theta_min, g.theta_max, groups=cfg0.groups())
        eng = Engine(self.lib, g, tab, device=self.device, n_spots=len(chunk))
        try:
            spots = [sp for (_, _, _, _, sp) in chunk]
            eng.run_batch(spots, tp)
            for s, (run, cfg, _, _, sp) in enumerate(chunk):
                edep = eng.batch_spot_edep(s)
                _, _, st = self.lib.fe_precompute(sp.E0, sp.sigma_x, g.Nz, g.dz)
                edep = self.lib.post_convolve(edep, g.dx, st, sp.sigma_x)
                out = run["config"].sections.get("[output]", {}).get("output_dir", run["dir"])
                if cfg.save_2d:
                    sio.write_dose_2d(self.lib, os.path.join(out, cfg.dose_2d_file.decode()), edep,
                                      g.dx, g.dz, bool(cfg.normalize_dose))
                if cfg.save_pdd:
                    sio.write_pdd(self.lib, os.path.join(out, cfg.pdd_file.decode()), edep, g.dz,
                                  bool(cfg.normalize_dose))
        finally:
            eng.close()

    def get_completed_dirs(self) -> list[str]:
        return [run["dir"] for run in self.completed]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="SM_2D batch sweep over sm2d_run_batch")
    ap.add_argument("-c", "--config", default="batch_config.yaml")
    ap.add_argument("--dry-run", action="store_true", help="list the runs and exit")
    ap.add_argument("--max-spots", type=int, default=MAX_SPOTS_PER_LAUNCH,
                    help="spots per batched launch")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    runner = BatchRunner(a.config, device=a.device)
    runs = runner.prepare_runs()
    if a.dry_run:
        print("runs:")
        for run in runs:
            print(f"  {run['name']}")
            for mod in run["params"]["modifications"]:
                print(f"    {mod['params']}")
        return 0
    runner.run_all(max_spots=a.max_spots)
    return 0 if not runner.failed else 1


if __name__ == "__main__":
    sys.exit(main())
