"""Writes the golden fixtures under tests/golden/ from the reference itself.

Run in the container that holds the reference sources (they never travel):

    make -C oracle ref            # builds oracle/_ref/ from /root/reference
    python tests/golden/make_golden.py

For each fixture: a small seeded index (written by this project's builder in
the reference's file format, committed as data), its queries, the scale and
code checksums of the reference's Index::load, and the ids / distances of the
reference's Index::search (AVX-512 engine) over a grid of parameters.
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "katoann-public_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from katoann import builder, datasets, formats  # noqa: E402
import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# name: (n, nq, d, nlist, centres, seed)
FIXTURES = {
    "g100": (4000, 128, 100, 16, 24, 101),   # d % 16 = 4: FMA tail of the 16-lane chains
    "g66": (3000, 96, 66, 12, 20, 102),     # d % 4 = 2: the plain-loop basis rows
}


def grid(nlist):
    out = []
    for ranking in ("prefix", "high"):
        for selection in ("local", "threshold"):
            for nprobe in (1, 4, min(16, nlist)):
                for k, S in ((10, 10), (10, 32), (10, 128), (32, 32)):
                    out.append((k, nprobe, S, ranking, selection))
    # Ranking::kFull keeps the k nearest of all candidates (survivors unused)
    for nprobe in (1, 4, min(16, nlist)):
        for k in (10, 32):
            out.append((k, nprobe, k, "full", "threshold"))
    return out


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    for name, (n, nq, d, nlist, centres, seed) in FIXTURES.items():
        base, q = datasets.mixture(n, nq, d=d, centres=centres, sigma=0.3, seed=seed)
        out_dir = os.path.join(HERE, name)
        if os.path.isdir(out_dir):
            shutil.rmtree(out_dir)
        os.makedirs(out_dir)
        prefix = os.path.join(out_dir, "idx")
        builder.build_numpy(base, prefix, nlist=nlist, kmeans_iters=6, seed=seed, normalized=True)
        formats.write_fbin(os.path.join(out_dir, "queries.fbin"), q)
        dump = os.path.join("/tmp", name + ".dump")
        subprocess.run([O.REF_SEARCH, "dump", prefix, dump], check=True)
        R = O.read_ref_dump(dump)
        codes = ((R["hi"].astype(np.uint16) << 8) | R["lo_plane"]).astype(np.uint16)
        ccodes = ((R["cen_hi"].astype(np.uint16) << 8) | R["cen_lo"]).astype(np.uint16)
        rot = os.path.join("/tmp", name + ".rot")
        subprocess.run([O.REF_SEARCH, "rotate", prefix, rot], check=True)
        rotated = np.fromfile(rot, np.float32)
        res = dict(lo=np.float32(R["lo"]), inv_step=np.float64(R["inv_step"]), step_sq=np.float64(R["step_sq"]),
                   prefix_dims=np.uint32(R["prefix_dims"]), codes_sha=sha(codes), ccodes_sha=sha(ccodes),
                   norm_sha=sha(R["norm"]), cnorm_sha=sha(R["cen_norm"]), rotated_sha=sha(rotated))
        cases = grid(nlist)
        for (k, nprobe, S, ranking, selection) in cases:
            ids, dists = O.ref_search(prefix, os.path.join(out_dir, "queries.fbin"), k, nprobe, S, ranking,
                                      selection)
            tag = f"{k}_{nprobe}_{S}_{ranking}_{selection}"
            res["ids_" + tag] = ids.reshape(nq, k)
            res["dists_" + tag] = dists.reshape(nq, k)
        np.savez_compressed(os.path.join(out_dir, "reference_outputs.npz"), **res)
        print(name, "cases", len(cases), "files", sorted(os.listdir(out_dir)))


if __name__ == "__main__":
    main()
