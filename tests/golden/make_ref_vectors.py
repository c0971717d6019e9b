"""Writes tests/golden/ref_vectors.json and tests/golden/mappings/*.so from the
reference's own code compiled here (oracle/_ref/libref.so, built by
oracle/Makefile from the reference tree).

Run in the build container, where the reference tree exists:
    make -C oracle _ref/libref.so && OMP_NUM_THREADS=8 python tests/golden/make_ref_vectors.py

The vectors are data only: sha256 digests of the reference's outputs (edge
lists, weights, relabelled CSRs, mappings), the Tier-0 feature values and the
modeled property bytes, on inputs every test can rebuild (the oracle's
generator and CSR build, seeded numpy integer draws).  The mapping files
are the reference's CPU reorderers on Kronecker scale 16 (edge factor 16):
RCM in its "bnf" variant (-o 11:bnf), the GraphBrew Leiden-Gorder8
composition (-o 12:leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8)
and the two adaptive rules' Leiden-Gorder arms, as binary int32 MAP files
(new_ids[old] = new), the format -o 13 reads; the drivers find them by name
under GBK_MAPPING_DIR.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402

MAPS = os.path.join(HERE, "mappings")
OUT = os.path.join(HERE, "ref_vectors.json")


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        if a is not None:
            h.update(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes())
    return h.hexdigest()


def directed_graph(weighted: bool):
    """A directed graph from seeded numpy integer draws (PCG64 integers are
    stable across numpy versions)."""
    rng = np.random.default_rng(5)
    n, m = 3000, 30000
    u = rng.integers(0, n, m).astype(np.int32)
    v = (rng.integers(0, n, m) * rng.integers(0, n, m) // n).astype(np.int32)  # skewed heads
    w = rng.integers(1, 256, m).astype(np.int32) if weighted else None
    return O.make_graph(u, v, w, symmetrize=False, n=n)


GRAPHS = {
    "kron12": lambda: O.kron(12),
    "uniform12": lambda: O.kron(12, uniform=True),
    "kron11_weighted": lambda: O.kron(11, weighted=True),
    "directed3000": lambda: directed_graph(False),
    "directed3000_weighted": lambda: directed_graph(True),
}


def main() -> None:
    if not R.available():
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle _ref/libref.so)")
    vec: dict = {"threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}

    gen = {}
    for scale in range(10, 17):
        for uniform in (False, True):
            u, v = R.generate_el(scale, 16, uniform)
            gen[f"{'u' if uniform else 'g'}{scale}"] = {
                "m": int(len(u)), "sha256": sha(u, v),
                "head": [[int(a), int(b)] for a, b in zip(u[:4], v[:4])]}
    vec["generate_el_ef16"] = gen
    m = (1 << 20) + 123
    w = R.insert_weights(m)
    vec["insert_weights"] = {"m": m, "sha256": sha(w), "head": [int(x) for x in w[:8]]}

    relabel, mappings, tier0 = {}, {}, {}
    for name, make in GRAPHS.items():
        g = make()
        ids = O.mapping_random(g.n)
        out = R.relabel(g, ids)
        relabel[name] = {"n": g.n, "mapping": "random", "sha256": sha(*out)}
        mp = {}
        for algo in (0, 1, 2, 3, 4, 5, 6, 7, 11):
            if algo == 11 and g.directed:
                continue
            mp[str(algo)] = sha(R.mapping(g, algo))
        mappings[name] = mp
        if not g.directed and g.wgt is None:
            tier0[name] = {"default": [float(x) for x in R.tier0_features(g)],
                           "s2000_seed7": [float(x) for x in R.tier0_features(g, 2000, 7)]}
    g14 = O.kron(14)
    tier0["kron14"] = {"default": [float(x) for x in R.tier0_features(g14)],
                       "s2000_seed7": [float(x) for x in R.tier0_features(g14, 2000, 7)]}
    vec["relabel"] = relabel
    vec["mapping_sha256"] = mappings
    vec["tier0"] = tier0
    vec["modeled_property_bytes"] = {
        str(b): R.modeled_property_bytes(b, 4194304, 134217728) for b in range(0, 9)}

    os.makedirs(MAPS, exist_ok=True)
    g16 = O.kron(16)
    files = {}
    arms = {
        # the all-kernel low-reuse rule's arm and the budgeted rule's arm
        # (reorder_adaptive.h:344-371, 262-283): parallel community
        # detection, so a run's output is one valid mapping among several
        "leiden_gorder8_lowreuse": R.GRAPHBREW_LEIDEN_GORDER8 + ":cd_parallel:sgmb4096:gordf5000:norefine:2:2",
        "leiden_gorder32_budgeted": "leiden:compose:sg_none:comm_identity:intra_gorder:gw32:gordf500:cd_parallel:1:1",
    }
    made = [("rcm_bnf", R.mapping(g16, 11)), ("leiden_gorder8", R.graphbrew_mapping(g16))]
    made += [(tag, R.graphbrew_mapping(g16, spec)) for tag, spec in arms.items()]
    for tag, ids in made:
        assert O.validate_permutation(ids) == 0
        path = os.path.join(MAPS, f"kron16_ef16_{tag}.so")
        ids.astype("<i4").tofile(path)
        files[tag] = {"file": os.path.relpath(path, HERE), "n": int(g16.n), "sha256": sha(ids),
                      "fingerprint": O.mapping_fingerprint(ids)}
    vec["mapping_files"] = files
    with open(OUT, "w") as f:
        json.dump(vec, f, indent=1, sort_keys=True)
        f.write("\n")
    print(OUT)


if __name__ == "__main__":
    main()
