"""Writes tests/golden/hostgen_golden.json from the REFERENCE's own host generator.

Runs oracle/_ref/ref_hostgen (the reference's config loader, grid-location, weight and
stimulus-schedule generators compiled from its sources by `make -C oracle ref`) on each config
below and records, per config: the parsed Config (every field), and for each of the four dense
weight matrices its shape, number of nonzeros, SHA-256 of the float32 bytes and SHA-256 of the
0/1 connection mask; the stimulus event arrays (and a digest of the per-step count table).

Run in the development container (the reference checkout is not on the GPU machines):
    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "ref_hostgen")
REF_NOFMA = os.path.join(REPO, "oracle", "_ref", "ref_hostgen_nofma")

CONFIGS = [
    ("config01", os.path.join(HERE, "configs", "config01.json")),
    ("defaults", os.path.join(HERE, "configs", "defaults.json")),
    ("mixed_keys", os.path.join(HERE, "configs", "mixed_keys.json")),
    ("small_offgrid", os.path.join(HERE, "configs", "small_offgrid.json")),
    ("net3125_stdp", os.path.join(REPO, "configs", "net3125_stdp.json")),
    ("net10125_stdp", os.path.join(REPO, "configs", "net10125_stdp.json")),
]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(binary: str, cfg: str, text: bool) -> dict:
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "ref_files"))
        subprocess.run([binary, cfg, d] + (["text"] if text else []), check=True)
        dump = json.load(open(os.path.join(d, "config_dump.json")))
        ne, ni = dump["num_excitatory_neurons"], dump["num_inhibitory_neurons"]
        shapes = [(ne, ne), (ne, ni), (ni, ne), (ni, ni)]
        mats = []
        for k in range(4):
            w = np.fromfile(os.path.join(d, f"w{k}.bin"), np.float32).reshape(shapes[k])
            mats.append({"shape": list(shapes[k]), "nnz": int((w != 0).sum()), "sha256": sha(w),
                         "mask_sha256": sha((w != 0).astype(np.uint8))})
        counts = np.fromfile(os.path.join(d, "stim_counts.bin"), np.uint32)
        stim = {
            "num_timesteps": int(counts.size),
            "counts_sha256": sha(counts),
            "nonzero_steps": np.nonzero(counts)[0].tolist()[:200],
            "neuron_ids": np.fromfile(os.path.join(d, "stim_ids.bin"), np.uint32).tolist()[:2000],
            "post_group_offsets": np.fromfile(os.path.join(d, "stim_offsets.bin"), np.uint32).tolist()[:2000],
            "amplitudes": np.fromfile(os.path.join(d, "stim_amps.bin"), np.float32).tolist()[:2000],
            "num_events": int(os.path.getsize(os.path.join(d, "stim_ids.bin")) // 4),
        }
        texts = {}
        if text:
            rf = os.path.join(d, "ref_files")
            for name in ("group_e_locations.txt", "group_i_locations.txt", "weights_e_to_i.txt"):
                with open(os.path.join(rf, name), "rb") as f:
                    texts[name] = hashlib.sha256(f.read()).hexdigest()
            sched = os.path.join(rf, "stimulus_schedules")
            for name in sorted(os.listdir(sched)):
                with open(os.path.join(sched, name), "rb") as f:
                    texts["stimulus_schedules/" + name] = hashlib.sha256(f.read()).hexdigest()
        return {"config": dump, "weights": mats, "stimulus": stim, "text_sha256": texts}


def main() -> None:
    if not os.path.exists(REF):
        sys.exit("build the reference generator first: make -C oracle ref")
    out = {}
    for name, path in CONFIGS:
        print("reference generator:", name, flush=True)
        small = name in ("defaults", "mixed_keys", "small_offgrid")
        rec = run(REF, path, text=small)
        rec["weights_nofma"] = [m["sha256"] for m in run(REF_NOFMA, path, text=False)["weights"]]
        out[name] = rec
    with open(os.path.join(HERE, "hostgen_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "hostgen_golden.json"))


if __name__ == "__main__":
    main()
