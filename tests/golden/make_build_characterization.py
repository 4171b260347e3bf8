"""Writes tests/golden/build_characterization.json from the reference's own
characterization record (tests/golden/build_characterization.txt, written by
tests/test_build_characterization.cpp:65-117 with the seed forced to 12345) and
the config files it covers.

For every config whose system the GPU path holds (alea_amd.config.build_spec
accepts it), the fixture keeps:
  * the config's `system`, `wavefunction`, `hamiltonian` and the methods and
    sample counts of its `qmc` stages, as parsed from the YAML (data only:
    comments and run settings are dropped);
  * the record's component lines and its four walker lines
    (log|psi|, E_L, Im E_L).
Configs it refuses are listed with the reason.  Nothing of the reference is
run or imported: both files are read as text.

Run in a tree that has /root/reference:
    python tests/golden/make_build_characterization.py
"""
import json
import os
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
from alea_amd import config  # noqa: E402


def read_record(path):
    out, cur = {}, None
    with open(path, encoding="utf-8") as f:
        for line in f:
            line = line.rstrip("\n")
            if line.startswith("config "):
                cur = line[7:]
                out[cur] = []
            elif cur is not None:
                out[cur].append(line)
    return out


def main():
    record = read_record(os.path.join(REF, "tests", "golden", "build_characterization.txt"))
    cases, refused = {}, {}
    for name, lines in sorted(record.items()):
        path = os.path.join(REF, name)
        if not os.path.exists(path) or not lines or lines[0].startswith(("error", "no-drivers")):
            refused[name] = "no walker record"
            continue
        with open(path, encoding="utf-8") as f:
            cfg = yaml.safe_load(f)
        data = {k: cfg.get(k) for k in ("system", "wavefunction", "hamiltonian")}
        data["qmc"] = [{"method": s.get("method"), "samples": s.get("samples")} for s in cfg.get("qmc", []) or []]
        if cfg.get("action"):
            data["action"] = True
        try:
            config.build_spec(data, requests=[])
            config.initial_positions(data, 1, 1)
        except ValueError as e:
            refused[name] = str(e)
            continue
        comps = [ln for ln in lines[1:] if not ln.startswith("  w")]
        walkers = [[float(x) for x in ln.split()[1:4]] for ln in lines if ln.startswith("  w")]
        cases[name] = dict(config=data, components=comps, walkers=walkers)
    out = dict(source="tests/golden/build_characterization.txt and the config files it names, "
                      "seed 12345, four walkers per config",
               seed=12345, cases=cases, refused=refused)
    with open(os.path.join(HERE, "build_characterization.json"), "w", encoding="utf-8") as f:
        json.dump(out, f, indent=1, ensure_ascii=False)
    print(f"{len(cases)} configs kept, {len(refused)} refused")


if __name__ == "__main__":
    main()
